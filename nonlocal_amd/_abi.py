"""ctypes mirror of the plain-data types declared in include/nlf.h."""
import ctypes as C
import math

NLF_OK = 0
NLF_ERR_ARGUMENT, NLF_ERR_HIP, NLF_ERR_STATE, NLF_ERR_UNSUPPORTED, NLF_ERR_NUMERIC = -1, -2, -3, -4, -5
NLF_ELEMENT_Q4 = 4
NLF_ELEMENT_Q8 = 8
NLF_INFLUENCE_POLYNOMIAL = 0
NLF_INFLUENCE_CONSTANT = 1
NLF_INFLUENCE_EXPONENTIAL = 2
NLF_DISTANCE_LP, NLF_DISTANCE_ELLIPSE_WITH_ROTATION = 0, 1
NLF_DISTANCE_N_INFINITY = -1
NLF_EXPAND_NO, NLF_EXPAND_MAX, NLF_EXPAND_MIN, NLF_EXPAND_MEAN = 0, 1, 2, 3
NLF_PART_LOCAL, NLF_PART_FULL = 0, 2
NLF_CONDUCTIVITY_ISOTROPIC, NLF_CONDUCTIVITY_ORTHOTROPIC, NLF_CONDUCTIVITY_ANISOTROPIC = 0, 1, 2
NLF_BC_NONE, NLF_BC_FIRST_KIND, NLF_BC_FLUX, NLF_BC_CONVECTION = 0, 1, 2, 3
NLF_ELASTIC_ISOTROPIC, NLF_ELASTIC_ORTHOTROPIC, NLF_ELASTIC_ANISOTROPIC = 0, 1, 2
NLF_EXPANSION_NONE, NLF_EXPANSION_ISOTROPIC, NLF_EXPANSION_ORTHOTROPIC, NLF_EXPANSION_ANISOTROPIC = 0, 1, 2, 3
NLF_PRECOND_IDENTITY, NLF_PRECOND_JACOBI, NLF_PRECOND_ILLT, NLF_PRECOND_LEARNED = 0, 1, 2, 3
NLF_MODE_FP64, NLF_MODE_MIXED, NLF_MODE_SPLIT = 0, 1, 2
NLF_OPT_ASSEMBLY_PAIRS = 1
NLF_OPT_ROW_PATTERNS = 2
NLF_OPT_ASSEMBLY_THREADS = 3
NLF_OPT_PATTERN_THREADS = 4
NLF_OPT_PAIR_BUDGET_MIB = 5
NLF_OPT_SPMV_SYMMETRIC = 6


class Model(C.Structure):
    _fields_ = [("local_weight", C.c_double), ("radius", C.c_double * 2), ("search_radius", C.c_double * 2),
                ("influence", C.c_int32), ("distance_n", C.c_int32), ("p", C.c_double), ("q", C.c_double),
                ("distance", C.c_int32), ("pad_", C.c_int32), ("distance_real", C.c_double)]


class Distance(C.Structure):
    """nlf_distance: the distance function of one group's neighbour search."""
    _fields_ = [("kind", C.c_int32), ("n", C.c_int32), ("n_real", C.c_double)]


def distance_exponent(n):
    """(distance_n, distance_real) of an exponent held like the reference's size_t_or<T>
    (read_influence_parameter, read_model.hpp:97-114): an int is the integer alternative, a float the
    real one, math.inf or "inf" / "infinity" the maximum norm."""
    if isinstance(n, str):
        if n in ("inf", "infinity"):
            return NLF_DISTANCE_N_INFINITY, 0.0
        raise ValueError(f"distance exponent {n!r} shall be a number or \"infinity\"")
    if isinstance(n, bool):
        raise ValueError("distance exponent shall be a number")
    if isinstance(n, float) and math.isinf(n):
        return NLF_DISTANCE_N_INFINITY, 0.0
    if isinstance(n, int) or type(n).__name__.startswith(("int", "uint")):
        if int(n) <= 0:
            raise ValueError("distance exponent shall be greater than 0")
        return int(n), 0.0
    if not float(n) > 0.0:
        raise ValueError("distance exponent shall be greater than 0")
    return 0, float(n)


def distance(kind=NLF_DISTANCE_LP, n=2):
    """nlf_distance of a kind (NLF_DISTANCE_*) and exponent n (ignored by the rotated ellipse)."""
    d = Distance()
    d.kind = kind
    d.n, d.n_real = distance_exponent(n) if kind == NLF_DISTANCE_LP else (2, 0.0)
    return d


def distance_of(m):
    """The nlf_distance of an nlf_model."""
    d = Distance()
    d.kind, d.n, d.n_real = m.distance, m.distance_n, m.distance_real
    return d


class ThermalMaterial(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad_", C.c_int32), ("c", C.c_double * 3)]


class ElasticMaterial(C.Structure):
    _fields_ = [("kind", C.c_int32), ("expansion_kind", C.c_int32), ("young", C.c_double), ("poisson", C.c_double),
                ("young_y", C.c_double), ("poisson_y", C.c_double), ("shear", C.c_double), ("angle", C.c_double),
                ("expansion", C.c_double * 3)]


class AssembleOptions(C.Structure):
    """nlf_assemble_options: the Neumann border and coefficients per quadrature point."""
    _fields_ = [("neumann", C.c_int32), ("pad_", C.c_int32), ("coefficients", C.POINTER(C.POINTER(C.c_double)))]


class Boundary(C.Structure):
    _fields_ = [("kind", C.c_int32 * 2), ("value", C.c_double * 2), ("heat_transfer", C.c_double),
                ("ambient", C.c_double)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int64), ("outer_iterations", C.c_int64), ("residual", C.c_double),
                ("true_residual", C.c_double), ("preconditioner", C.c_int32), ("fallbacks", C.c_int32),
                ("seconds", C.c_double), ("spmv_ms", C.c_double), ("precond_ms", C.c_double),
                ("spmv_samples", C.c_int64), ("spmv_bytes", C.c_int64)]


def model(local_weight, radius, search_radius=None, influence=NLF_INFLUENCE_POLYNOMIAL, p=2.0, q=1.0,
          distance=NLF_DISTANCE_LP, n=2):
    """Build an nlf_model; `radius` may be a scalar (isotropic) or a pair; `distance` NLF_DISTANCE_* with
    the exponent `n` of the Lp distance (see distance_exponent)."""
    m = Model()
    m.local_weight = float(local_weight)
    r = (radius, radius) if not hasattr(radius, "__len__") else radius
    m.radius[0], m.radius[1] = float(r[0]), float(r[1])
    if search_radius is None:
        search_radius = r
    s = (search_radius, search_radius) if not hasattr(search_radius, "__len__") else search_radius
    m.search_radius[0], m.search_radius[1] = float(s[0]), float(s[1])
    m.influence = influence
    m.distance = distance
    m.distance_n, m.distance_real = distance_exponent(n) if distance == NLF_DISTANCE_LP else (2, 0.0)
    m.p = float(p)
    m.q = float(q)
    return m


def thermal_material(kind=NLF_CONDUCTIVITY_ISOTROPIC, c=(1.0, 0.0, 0.0)):
    t = ThermalMaterial()
    t.kind = kind
    for i, v in enumerate(c):
        t.c[i] = float(v)
    return t


def elastic_material(young=1.0, poisson=0.3, expansion=None):
    """Isotropic plane-stress material (isotropic_elastic_parameters); `expansion` an optional thermal
    expansion coefficient: a scalar (isotropic), a pair (orthotropic) or a triple (anisotropic)."""
    e = ElasticMaterial()
    e.kind = NLF_ELASTIC_ISOTROPIC
    e.young = float(young)
    e.poisson = float(poisson)
    _set_expansion(e, expansion)
    return e


def orthotropic_material(young, poisson, shear, expansion=None):
    """orthotropic_elastic_parameters: young = (E_x, E_y), poisson = (nu_0, nu_1), shear = G."""
    e = ElasticMaterial()
    e.kind = NLF_ELASTIC_ORTHOTROPIC
    e.young, e.young_y = float(young[0]), float(young[1])
    e.poisson, e.poisson_y = float(poisson[0]), float(poisson[1])
    e.shear = float(shear)
    _set_expansion(e, expansion)
    return e


def anisotropic_material(young, poisson, shear, angle, expansion=None):
    """anisotropic_elastic_parameters: the orthotropic material rotated by `angle` (radians)."""
    e = orthotropic_material(young, poisson, shear, expansion)
    e.kind = NLF_ELASTIC_ANISOTROPIC
    e.angle = float(angle)
    return e


def _set_expansion(e, expansion):
    if expansion is None:
        e.expansion_kind = NLF_EXPANSION_NONE
        return
    v = [expansion] if not hasattr(expansion, "__len__") else list(expansion)
    e.expansion_kind = {1: NLF_EXPANSION_ISOTROPIC, 2: NLF_EXPANSION_ORTHOTROPIC, 3: NLF_EXPANSION_ANISOTROPIC}[len(v)]
    for i, x in enumerate(v):
        e.expansion[i] = float(x)


def boundary(kind=NLF_BC_NONE, value=0.0, kind2=NLF_BC_NONE, value2=0.0, heat_transfer=0.0, ambient=0.0):
    b = Boundary()
    b.kind[0], b.kind[1] = kind, kind2
    b.value[0], b.value[1] = float(value), float(value2)
    b.heat_transfer = float(heat_transfer)
    b.ambient = float(ambient)
    return b


def array_of(ctype, items):
    arr = (ctype * len(items))()
    for i, it in enumerate(items):
        arr[i] = it
    return arr


# nlf_boundary_value_fn: double (*)(double x, double y, int group, int dof, void* user)
BoundaryValueFn = C.CFUNCTYPE(C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_void_p)


def boundary_callback(values):
    """ctypes callback of a Python callable values(x, y, group, dof) -> float (None -> None).  Keep the
    returned object alive while the library may call it."""
    if values is None:
        return None
    return BoundaryValueFn(lambda x, y, g, d, _user: float(values(x, y, g, d)))

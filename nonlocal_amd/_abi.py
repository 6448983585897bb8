"""ctypes mirror of the plain-data types declared in include/nlf.h."""
import ctypes as C

NLF_OK = 0
NLF_ERR_ARGUMENT, NLF_ERR_HIP, NLF_ERR_STATE, NLF_ERR_UNSUPPORTED, NLF_ERR_NUMERIC = -1, -2, -3, -4, -5
NLF_ELEMENT_Q4 = 4
NLF_ELEMENT_Q8 = 8
NLF_INFLUENCE_POLYNOMIAL = 0
NLF_INFLUENCE_CONSTANT = 1
NLF_INFLUENCE_EXPONENTIAL = 2
NLF_EXPAND_NO, NLF_EXPAND_MAX, NLF_EXPAND_MIN, NLF_EXPAND_MEAN = 0, 1, 2, 3
NLF_PART_LOCAL, NLF_PART_FULL = 0, 2
NLF_CONDUCTIVITY_ISOTROPIC, NLF_CONDUCTIVITY_ORTHOTROPIC, NLF_CONDUCTIVITY_ANISOTROPIC = 0, 1, 2
NLF_BC_NONE, NLF_BC_FIRST_KIND, NLF_BC_FLUX, NLF_BC_CONVECTION = 0, 1, 2, 3
NLF_PRECOND_IDENTITY, NLF_PRECOND_JACOBI, NLF_PRECOND_ILLT, NLF_PRECOND_LEARNED = 0, 1, 2, 3
NLF_MODE_FP64, NLF_MODE_MIXED, NLF_MODE_SPLIT = 0, 1, 2
NLF_OPT_ASSEMBLY_PAIRS = 1
NLF_OPT_ROW_PATTERNS = 2
NLF_OPT_ASSEMBLY_THREADS = 3
NLF_OPT_PATTERN_THREADS = 4
NLF_OPT_PAIR_BUDGET_MIB = 5


class Model(C.Structure):
    _fields_ = [("local_weight", C.c_double), ("radius", C.c_double * 2), ("search_radius", C.c_double * 2),
                ("influence", C.c_int32), ("distance_n", C.c_int32), ("p", C.c_double), ("q", C.c_double)]


class ThermalMaterial(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad_", C.c_int32), ("c", C.c_double * 3)]


class ElasticMaterial(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad_", C.c_int32), ("young", C.c_double), ("poisson", C.c_double)]


class Boundary(C.Structure):
    _fields_ = [("kind", C.c_int32 * 2), ("value", C.c_double * 2), ("heat_transfer", C.c_double),
                ("ambient", C.c_double)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int64), ("outer_iterations", C.c_int64), ("residual", C.c_double),
                ("true_residual", C.c_double), ("preconditioner", C.c_int32), ("fallbacks", C.c_int32),
                ("seconds", C.c_double), ("spmv_ms", C.c_double), ("precond_ms", C.c_double),
                ("spmv_samples", C.c_int64), ("spmv_bytes", C.c_int64)]


def model(local_weight, radius, search_radius=None, influence=NLF_INFLUENCE_POLYNOMIAL, p=2.0, q=1.0):
    """Build an nlf_model; `radius` may be a scalar (isotropic) or a pair."""
    m = Model()
    m.local_weight = float(local_weight)
    r = (radius, radius) if not hasattr(radius, "__len__") else radius
    m.radius[0], m.radius[1] = float(r[0]), float(r[1])
    if search_radius is None:
        search_radius = r
    s = (search_radius, search_radius) if not hasattr(search_radius, "__len__") else search_radius
    m.search_radius[0], m.search_radius[1] = float(s[0]), float(s[1])
    m.influence = influence
    m.distance_n = 2
    m.p = float(p)
    m.q = float(q)
    return m


def thermal_material(kind=NLF_CONDUCTIVITY_ISOTROPIC, c=(1.0, 0.0, 0.0)):
    t = ThermalMaterial()
    t.kind = kind
    for i, v in enumerate(c):
        t.c[i] = float(v)
    return t


def elastic_material(young=1.0, poisson=0.3):
    e = ElasticMaterial()
    e.kind = 0
    e.young = float(young)
    e.poisson = float(poisson)
    return e


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

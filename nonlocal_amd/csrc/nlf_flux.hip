// Post-processing of a solution: the heat flux (heat_equation_solution_2d::calc_flux,
// solvers/solver_2d/thermal/heat_equation_solution_2d.hpp:99-137) and the strains and stresses
// (mechanical_solution_2d::calc_strain_and_stress, solvers/solver_2d/mechanical/
// mechanical_solution_2d.hpp:192-232).  Both are "local value at every quadrature point, then its
// nonlocal average, then node values", with C = 2 (flux) or 3 (stress) components:
//
//   1. k_flux_local / k_stress_local (one thread per quadrature point): the gradient
//      sum_i u_i D_i(q) / |J_q| (gradient_in_qnodes, mesh_2d_utils.hpp:39-63), then
//      -lambda grad T (local_flux_in_qnodes, heat_equation_solution_2d.hpp:72-97) or the strain
//      (u_x,x, u_y,y, (u_x,y + u_y,x)/2) and the stress C eps (strains_in_quadratures,
//      mechanical_solution_2d.hpp:150-157; calc_stress, mechanical_parameters_2d.hpp:27-47).
//      Writes the local values [C][Q] and, per quadrature point, the record
//      {x_q, y_q, w_q |J_q| v_0, w_q |J_q| v_1} (+ w_q |J_q| v_2 for C = 3) that the gather streams.
//   2. k_nonlocal_average (one wave per element eL): for a nonlocal group
//        v(x_L) = p1 v_loc(x_L) + (1 - p1) sum_{eNL in N(eL)} sum_qNL w |J| phi(x_L, x_NL) v_loc(x_NL)
//      (heat_equation_solution_2d.hpp:106-131; calc_nonlocal_stress, mechanical_solution_2d.hpp:
//      173-189).  The lanes run over the (neighbour, quadrature point) pairs of eL, so consecutive
//      lanes read consecutive records, and each lane evaluates phi against the nq points of eL; a
//      wave reduction gives the nq sums.  Local groups copy v_loc.
//   3. k_to_nodes (one thread per node): qnodes_to_nodes (mesh_2d_utils.hpp:82-101), the
//      area-weighted mean over the node's elements (ascending ids) of the value at the element's
//      quadrature point nearest to the node; mesh_2d::area (mesh_2d.hpp:226-235).
//
// The sums run in a different order than the reference's loops, so results agree with it to
// rounding, not bit for bit.
#include "nlf_internal.hpp"

#include <vector>

namespace nlf {
namespace {

template<int NN, int NQ>
__global__ __launch_bounds__(256) void k_flux_local(i64 E, const i32* __restrict__ el, const double* __restrict__ D,
                                                    const double* __restrict__ jac, const double* __restrict__ xq,
                                                    const i32* __restrict__ group_of,
                                                    const GroupParams* __restrict__ groups, FeTables t,
                                                    const double* __restrict__ T, double* __restrict__ qloc,
                                                    double4* __restrict__ rec) {
    const i64 idx = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    const i64 Q = E * NQ;
    if (idx >= Q) return;
    const i64 e = idx / NQ;
    const int q = int(idx - e * NQ);
    double gx = 0.0, gy = 0.0;
    for (int i = 0; i < NN; ++i) {
        const double v = T[el[e * NN + i]];
        const i64 d = 2 * ((e * NN + i) * NQ + q);
        gx += v * D[d];
        gy += v * D[d + 1];
    }
    const double j = jac[idx];
    gx /= j;
    gy /= j;
    const GroupParams& g = groups[group_of[e]];
    double fx, fy;
    if (g.ckind == NLF_CONDUCTIVITY_ISOTROPIC) { fx = -g.c[0] * gx; fy = -g.c[0] * gy; }
    else if (g.ckind == NLF_CONDUCTIVITY_ORTHOTROPIC) { fx = -g.c[0] * gx; fy = -g.c[1] * gy; }
    else { fx = -g.c[0] * gx - g.c[2] * gy; fy = -g.c[2] * gx - g.c[1] * gy; }
    qloc[idx] = fx;
    qloc[Q + idx] = fy;
    const double wj = t.w[q] * j;
    rec[idx] = make_double4(xq[2 * idx], xq[2 * idx + 1], wj * fx, wj * fy);
}

// calc_stress (mechanical_parameters_2d.hpp:27-48) with the Hooke matrix in the anisotropic layout
// {h11, h12, h66, h22, h16, h26} (isotropic and orthotropic: h16 = h26 = 0); tensor shear strain.
__device__ __forceinline__ void hooke_stress(const GroupParams& g, double e11, double e22, double e12, double& s11,
                                             double& s22, double& s12) {
    s11 = g.c[0] * e11 + g.c[1] * e22 + 2 * g.c[4] * e12;
    s22 = g.c[1] * e11 + g.c[3] * e22 + 2 * g.c[5] * e12;
    s12 = g.c[4] * e11 + g.c[5] * e22 + 2 * g.c[2] * e12;
}

// temperature change at quadrature point q of element e: sum_i N_i(q) dT_i (nodes_to_qnodes,
// mesh_2d_utils.hpp:65-80)
template<int NN, int NQ>
__device__ __forceinline__ double temperature_at(const i32* __restrict__ el, const FeTables& t,
                                                 const double* __restrict__ dT, i64 e, int q) {
    double v = 0.0;
    for (int i = 0; i < NN; ++i) v += dT[el[e * NN + i]] * t.qN[i * NQ + q];
    return v;
}

// u is the solution vector (x, y per node).  The strains written to qstrain are those of u; with a
// temperature change dT (node values, or null) the thermal strain alpha dT of the group is
// subtracted before the stress (substract_temperature_strains, mechanical_solution_2d.hpp:135-166).
template<int NN, int NQ>
__global__ __launch_bounds__(256) void k_stress_local(i64 E, const i32* __restrict__ el, const double* __restrict__ D,
                                                      const double* __restrict__ jac, const double* __restrict__ xq,
                                                      const i32* __restrict__ group_of,
                                                      const GroupParams* __restrict__ groups, FeTables t,
                                                      const double* __restrict__ u, const double* __restrict__ dT,
                                                      double* __restrict__ qstrain, double* __restrict__ qloc,
                                                      double4* __restrict__ rec, double* __restrict__ rec2) {
    const i64 idx = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    const i64 Q = E * NQ;
    if (idx >= Q) return;
    const i64 e = idx / NQ;
    const int q = int(idx - e * NQ);
    double g11 = 0.0, g12 = 0.0, g21 = 0.0, g22 = 0.0;       // d u_x / dx, d u_x / dy, d u_y / dx, d u_y / dy
    for (int i = 0; i < NN; ++i) {
        const i64 node = el[e * NN + i];
        const double ux = u[2 * node], uy = u[2 * node + 1];
        const i64 d = 2 * ((e * NN + i) * NQ + q);
        const double dx = D[d], dy = D[d + 1];
        g11 += ux * dx;
        g12 += ux * dy;
        g21 += uy * dx;
        g22 += uy * dy;
    }
    const double j = jac[idx];
    g11 /= j;
    g12 /= j;
    g21 /= j;
    g22 /= j;
    double e11 = g11, e22 = g22, e12 = 0.5 * (g12 + g21);
    qstrain[idx] = e11;
    qstrain[Q + idx] = e22;
    qstrain[2 * Q + idx] = e12;
    const GroupParams& g = groups[group_of[e]];
    if (dT && g.ekind != NLF_EXPANSION_NONE) {
        const double dt = temperature_at<NN, NQ>(el, t, dT, e, q);
        e11 -= g.alpha[0] * dt;
        e22 -= g.alpha[1] * dt;
        e12 -= g.alpha[2] * dt;
    }
    double s11, s22, s12;
    hooke_stress(g, e11, e22, e12, s11, s22, s12);
    qloc[idx] = s11;
    qloc[Q + idx] = s22;
    qloc[2 * Q + idx] = s12;
    const double wj = t.w[q] * j;
    rec[idx] = make_double4(xq[2 * idx], xq[2 * idx + 1], wj * s11, wj * s22);
    rec2[idx] = wj * s12;
}

// The stress of the thermal strain at each quadrature point (_temperature_condition::calc_stress,
// temperature_condition_2d.hpp:20-74: C (alpha dT), the shear part with the tensor convention), in
// the layout of k_stress_local, so that k_nonlocal_average forms its nonlocal average.
template<int NN, int NQ>
__global__ __launch_bounds__(256) void k_thermal_stress(i64 E, const i32* __restrict__ el,
                                                        const double* __restrict__ jac, const double* __restrict__ xq,
                                                        const i32* __restrict__ group_of,
                                                        const GroupParams* __restrict__ groups, FeTables t,
                                                        const double* __restrict__ dT, double* __restrict__ qloc,
                                                        double4* __restrict__ rec, double* __restrict__ rec2) {
    const i64 idx = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    const i64 Q = E * NQ;
    if (idx >= Q) return;
    const i64 e = idx / NQ;
    const int q = int(idx - e * NQ);
    const GroupParams& g = groups[group_of[e]];
    double s11 = 0.0, s22 = 0.0, s12 = 0.0;
    if (g.ekind != NLF_EXPANSION_NONE) {
        const double dt = temperature_at<NN, NQ>(el, t, dT, e, q);
        hooke_stress(g, g.alpha[0] * dt, g.alpha[1] * dt, g.alpha[2] * dt, s11, s22, s12);
    }
    qloc[idx] = s11;
    qloc[Q + idx] = s22;
    qloc[2 * Q + idx] = s12;
    const double wj = t.w[q] * jac[idx];
    rec[idx] = make_double4(xq[2 * idx], xq[2 * idx + 1], wj * s11, wj * s22);
    rec2[idx] = wj * s12;
}

// temperature_condition (temperature_condition_2d.hpp:128-170), one thread per node: over the node's
// elements (ascending), p1 times the local load sum_q w_q D_i(q) . s(q) (:80-92) plus, for a nonlocal
// group, (1 - p1) times sum_q w_q (D_i,x (S_xx + S_xy), D_i,y (S_yy + S_xy)) with S the influence
// average of the thermal stress over the neighbours -- the reference's nonlocal form (:118-121).
// qs holds s, qavg holds S (k_nonlocal_average with p1 = 0).
template<int NN, int NQ>
__global__ __launch_bounds__(256) void k_thermal_load(i64 N, const i32* __restrict__ ne_off,
                                                      const i32* __restrict__ ne_ids, const i32* __restrict__ el,
                                                      const double* __restrict__ D, const i32* __restrict__ group_of,
                                                      const GroupParams* __restrict__ groups, FeTables t, i64 Q,
                                                      const double* __restrict__ qs, const double* __restrict__ qavg,
                                                      double* __restrict__ f) {
    const i64 node = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (node >= N) return;
    double fx = 0.0, fy = 0.0;
    for (i32 k = ne_off[node]; k < ne_off[node + 1]; ++k) {
        const i64 e = ne_ids[k];
        const GroupParams& g = groups[group_of[e]];
        if (g.ekind == NLF_EXPANSION_NONE) continue;
        int i = 0;
        for (int j = 0; j < NN; ++j)
            if (el[e * NN + j] == node) i = j;
        double lx = 0.0, ly = 0.0, nx = 0.0, ny = 0.0;
        for (int q = 0; q < NQ; ++q) {
            const i64 d = 2 * ((e * NN + i) * NQ + q);
            const double wx = t.w[q] * D[d], wy = t.w[q] * D[d + 1];
            const i64 at = e * NQ + q;
            lx += wx * qs[at] + wy * qs[2 * Q + at];
            ly += wy * qs[Q + at] + wx * qs[2 * Q + at];
            if (g.nonlocal) {
                nx += wx * (qavg[at] + qavg[2 * Q + at]);
                ny += wy * (qavg[Q + at] + qavg[2 * Q + at]);
            }
        }
        if (g.nonlocal) {
            fx += (1.0 - g.p1) * nx;
            fy += (1.0 - g.p1) * ny;
        }
        fx += g.p1 * lx;
        fy += g.p1 * ly;
    }
    f[2 * node] += fx;
    f[2 * node + 1] += fy;
}

template<int NQ, int POLY, int C>
__global__ __launch_bounds__(256) void k_nonlocal_average(i64 E, const i64* __restrict__ nb_off,
                                                          const i32* __restrict__ nb_ids,
                                                          const i32* __restrict__ group_of,
                                                          const GroupParams* __restrict__ groups,
                                                          const double4* __restrict__ rec,
                                                          const double* __restrict__ rec2,
                                                          const double* __restrict__ qloc, double* __restrict__ qout) {
    // eL is the same for the whole wave; saying so lets the group parameters, the neighbour range
    // and every branch on the influence kind live in scalar registers
    const i64 eL = __builtin_amdgcn_readfirstlane(int(blockIdx.x) * 4 + int(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    if (eL >= E) return;
    const i64 Q = E * NQ;
    const GroupParams& g = groups[__builtin_amdgcn_readfirstlane(group_of[eL])];
    if (!g.nonlocal) {
        if (lane < NQ)
            for (int c = 0; c < C; ++c) qout[c * Q + eL * NQ + lane] = qloc[c * Q + eL * NQ + lane];
        return;
    }
    const InfluenceParams inf = g.inf;
    double xl[NQ], yl[NQ], acc[C][NQ];
    for (int q = 0; q < NQ; ++q) {
        const double4 r = rec[eL * NQ + q];
        xl[q] = r.x;
        yl[q] = r.y;
        for (int c = 0; c < C; ++c) acc[c][q] = 0.0;
    }
    const i64 b = nb_off[eL];
    const i64 pairs = (nb_off[eL + 1] - b) * NQ;
    // pair t = (neighbour t / NQ, its quadrature point t % NQ).  Software pipeline: while pair t is
    // evaluated, the record of pair t + 64 and the neighbour id of pair t + 128 are in flight, so
    // the id -> record load chain is not waited for in the loop.
    i64 t = lane;
    double4 r = make_double4(0.0, 0.0, 0.0, 0.0);
    double r2 = 0.0;
    i32 id_next = 0;
    if (t < pairs) {
        const i64 s = i64(nb_ids[b + t / NQ]) * NQ + t % NQ;
        r = rec[s];
        if (C == 3) r2 = rec2[s];
    }
    if (t + 64 < pairs) id_next = nb_ids[b + (t + 64) / NQ];
    for (; t < pairs; t += 64) {
        double4 r_next = make_double4(0.0, 0.0, 0.0, 0.0);
        double r2_next = 0.0;
        i32 id_after = 0;
        if (t + 64 < pairs) {
            const i64 s = i64(id_next) * NQ + (t + 64) % NQ;
            r_next = rec[s];
            if (C == 3) r2_next = rec2[s];
        }
        if (t + 128 < pairs) id_after = nb_ids[b + (t + 128) / NQ];
        for (int q = 0; q < NQ; ++q) {
            const double phi = influence_poly<POLY>(inf, xl[q], yl[q], r.x, r.y);
            acc[0][q] += phi * r.z;
            acc[1][q] += phi * r.w;
            if (C == 3) acc[C - 1][q] += phi * r2;
        }
        r = r_next;
        r2 = r2_next;
        id_next = id_after;
    }
    for (int c = 0; c < C; ++c)
        for (int q = 0; q < NQ; ++q)
            for (int o = 32; o > 0; o >>= 1) acc[c][q] += __shfl_xor(acc[c][q], o);
    if (lane < NQ) {
        const double p1 = g.p1, w1 = 1.0 - g.p1;
        for (int c = 0; c < C; ++c) {
            double sum = 0.0;
            for (int q = 0; q < NQ; ++q)
                if (q == lane) sum = acc[c][q];
            const i64 at = c * Q + eL * NQ + lane;
            qout[at] = qloc[at] * p1 + sum * w1;
        }
    }
}

template<int NN, int NQ, int C>
__global__ __launch_bounds__(256) void k_to_nodes(i64 N, i64 Q, const i32* __restrict__ ne_off,
                                                  const i32* __restrict__ ne_ids, const i32* __restrict__ el,
                                                  const double* __restrict__ jac, FeTables t,
                                                  const double* __restrict__ qv, double* __restrict__ out) {
    const i64 node = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (node >= N) return;
    double v[C], node_area = 0.0;
    for (int c = 0; c < C; ++c) v[c] = 0.0;
    for (i32 k = ne_off[node]; k < ne_off[node + 1]; ++k) {
        const i64 e = ne_ids[k];
        double area = 0.0;
        for (int q = 0; q < NQ; ++q) {
            const double factor = t.w[q] * jac[e * NQ + q];
            for (int i = 0; i < NN; ++i) area += factor * t.qN[i * NQ + q];
        }
        int local = 0;
        for (int i = 0; i < NN; ++i)
            if (el[e * NN + i] == node) local = i;
        const i64 qs = e * NQ + t.nearest[local];
        for (int c = 0; c < C; ++c) v[c] += area * qv[c * Q + qs];
        node_area += area;
    }
    for (int c = 0; c < C; ++c) out[c * N + node] = v[c] / node_area;
}

template<int NQ, int C>
void launch_average(nlf_mesh m, const GroupParams* groups, int poly, const double4* rec, const double* rec2,
                    const double* qloc, double* qout, hipStream_t s) {
    auto kernel = poly == 2 ? k_nonlocal_average<NQ, 2, C>
                            : (poly == 1 ? k_nonlocal_average<NQ, 1, C> : k_nonlocal_average<NQ, 0, C>);
    hipLaunchKernelGGL(kernel, dim3(blocks_for(m->E, 4)), dim3(256), 0, s, m->E,
                       m->has_neighbours ? m->nb_off.p : nullptr, m->has_neighbours ? m->nb_ids.p : nullptr,
                       m->group_of.p, groups, rec, rec2, qloc, qout);
    NLF_HIP(hipGetLastError());
}

template<int NN, int NQ, int C>
void launch_to_nodes(nlf_mesh m, const double* qv, double* out, hipStream_t s) {
    hipLaunchKernelGGL((k_to_nodes<NN, NQ, C>), dim3(blocks_for(m->N, 256)), dim3(256), 0, s, m->N, m->E * NQ,
                       m->ne_off.p, m->ne_ids.p, m->el.p, m->jac.p, m->fe, qv, out);
    NLF_HIP(hipGetLastError());
}

template<int NN, int NQ>
void launch_flux(nlf_mesh m, const GroupParams* groups, int poly, const double* T, double* qloc, double4* rec,
                 double* qout, double* out, hipStream_t s) {
    hipLaunchKernelGGL((k_flux_local<NN, NQ>), dim3(blocks_for(m->E * NQ, 256)), dim3(256), 0, s, m->E, m->el.p,
                       m->D.p, m->jac.p, m->xq.p, m->group_of.p, groups, m->fe, T, qloc, rec);
    NLF_HIP(hipGetLastError());
    launch_average<NQ, 2>(m, groups, poly, rec, nullptr, qloc, qout, s);
    if (out) launch_to_nodes<NN, NQ, 2>(m, qout, out, s);
}

template<int NN, int NQ>
void launch_stress(nlf_mesh m, const GroupParams* groups, int poly, const double* u, const double* dT, double* qstrain,
                   double* qloc, double4* rec, double* rec2, double* qout, double* strain, double* stress,
                   hipStream_t s) {
    hipLaunchKernelGGL((k_stress_local<NN, NQ>), dim3(blocks_for(m->E * NQ, 256)), dim3(256), 0, s, m->E, m->el.p,
                       m->D.p, m->jac.p, m->xq.p, m->group_of.p, groups, m->fe, u, dT, qstrain, qloc, rec, rec2);
    NLF_HIP(hipGetLastError());
    if (strain) launch_to_nodes<NN, NQ, 3>(m, qstrain, strain, s);
    if (!qout) return;
    launch_average<NQ, 3>(m, groups, poly, rec, rec2, qloc, qout, s);
    if (stress) launch_to_nodes<NN, NQ, 3>(m, qout, stress, s);
}

// Influence class of the nonlocal groups for gather_phi: 2 (bell), 1 (fast polynomial) when all
// nonlocal groups share it, else 0 (general influence()); -1 when no group is nonlocal.
bool any_nonlocal(const std::vector<GroupParams>& gp) {
    for (const GroupParams& g : gp)
        if (g.nonlocal) return true;
    return false;
}

}  // namespace
}  // namespace nlf

using namespace nlf;

extern "C" int nlf_heat_flux(nlf_mesh m, const nlf_model* models, const nlf_thermal_material* materials,
                             const double* T_dev, double* flux_dev, double* qflux_dev) {
    if (!m || !models || !materials || !T_dev || (!flux_dev && !qflux_dev)) return NLF_ERR_ARGUMENT;
    return guarded(m->ctx, [&] {
        hipStream_t s = m->ctx->stream;
        std::vector<GroupParams> gp(size_t(m->G));
        for (int g = 0; g < m->G; ++g) {
            gp[g] = model_group_params(models[g], false);
            const nlf_thermal_material& t = materials[g];
            NLF_CHECK(t.kind >= NLF_CONDUCTIVITY_ISOTROPIC && t.kind <= NLF_CONDUCTIVITY_ANISOTROPIC, NLF_ERR_ARGUMENT,
                      "unknown conductivity kind");
            gp[g].ckind = t.kind;
            gp[g].c[0] = t.c[0]; gp[g].c[1] = t.c[1]; gp[g].c[2] = t.c[2];
        }
        NLF_CHECK(!any_nonlocal(gp) || m->has_neighbours, NLF_ERR_STATE,
                  "the flux of a nonlocal group needs the neighbour lists (nlf_find_neighbours)");
        if (m->E == 0) return;
        const i64 Q = m->E * m->fe.nq;
        DevBuf<GroupParams> d_groups;
        upload(d_groups, gp.data(), gp.size(), s);
        DevBuf<double> qloc{size_t(2 * Q)};
        DevBuf<double4> rec{size_t(Q)};
        DevBuf<double> own;
        double* qout = qflux_dev;
        if (!qout) {
            own.alloc(size_t(2 * Q));
            qout = own.p;
        }
        const int poly = influence_class(gp);
        if (m->nn == NLF_ELEMENT_Q4) launch_flux<4, 4>(m, d_groups.p, poly, T_dev, qloc.p, rec.p, qout, flux_dev, s);
        else launch_flux<8, 9>(m, d_groups.p, poly, T_dev, qloc.p, rec.p, qout, flux_dev, s);
        NLF_HIP(hipStreamSynchronize(s));
    });
}

extern "C" int nlf_strain_stress(nlf_mesh m, const nlf_model* models, const nlf_elastic_material* materials,
                                 const double* u_dev, const double* delta_temperature_dev, double* strain_dev,
                                 double* stress_dev, double* qstress_dev) {
    if (!m || !models || !materials || !u_dev || (!strain_dev && !stress_dev && !qstress_dev)) return NLF_ERR_ARGUMENT;
    return guarded(m->ctx, [&] {
        hipStream_t s = m->ctx->stream;
        std::vector<GroupParams> gp(size_t(m->G));
        for (int g = 0; g < m->G; ++g) {
            gp[g] = model_group_params(models[g], false);
            elastic_group_params(materials[g], gp[g]);
        }
        NLF_CHECK(!any_nonlocal(gp) || m->has_neighbours, NLF_ERR_STATE,
                  "the stress of a nonlocal group needs the neighbour lists (nlf_find_neighbours)");
        if (m->E == 0) return;
        const i64 Q = m->E * m->fe.nq;
        DevBuf<GroupParams> d_groups;
        upload(d_groups, gp.data(), gp.size(), s);
        DevBuf<double> qstrain{size_t(3 * Q)}, qloc{size_t(3 * Q)}, rec2{size_t(Q)};
        DevBuf<double4> rec{size_t(Q)};
        DevBuf<double> own;
        double* qout = qstress_dev;
        if (!qout && stress_dev) {
            own.alloc(size_t(3 * Q));
            qout = own.p;
        }
        const int poly = influence_class(gp);
        if (m->nn == NLF_ELEMENT_Q4)
            launch_stress<4, 4>(m, d_groups.p, poly, u_dev, delta_temperature_dev, qstrain.p, qloc.p, rec.p, rec2.p, qout,
                                strain_dev, stress_dev, s);
        else
            launch_stress<8, 9>(m, d_groups.p, poly, u_dev, delta_temperature_dev, qstrain.p, qloc.p, rec.p, rec2.p, qout,
                                strain_dev, stress_dev, s);
        NLF_HIP(hipStreamSynchronize(s));
    });
}

extern "C" int nlf_temperature_condition(nlf_mesh m, const nlf_model* models, const nlf_elastic_material* materials,
                                         const double* delta_temperature_dev, double* f_dev) {
    if (!m || !models || !materials || !delta_temperature_dev || !f_dev) return NLF_ERR_ARGUMENT;
    return guarded(m->ctx, [&] {
        hipStream_t s = m->ctx->stream;
        std::vector<GroupParams> gp(size_t(m->G)), avg(size_t(m->G));
        bool any = false;
        for (int g = 0; g < m->G; ++g) {
            gp[g] = model_group_params(models[g], false);
            elastic_group_params(materials[g], gp[g]);
            any |= gp[g].ekind != NLF_EXPANSION_NONE;
            avg[g] = gp[g];
            avg[g].p1 = 0.0;                     // k_nonlocal_average then returns the average alone
        }
        NLF_CHECK(!any_nonlocal(gp) || m->has_neighbours, NLF_ERR_STATE,
                  "the thermal load of a nonlocal group needs the neighbour lists (nlf_find_neighbours)");
        if (m->E == 0 || !any) return;
        const i64 Q = m->E * m->fe.nq;
        DevBuf<GroupParams> d_groups, d_avg;
        upload(d_groups, gp.data(), gp.size(), s);
        upload(d_avg, avg.data(), avg.size(), s);
        DevBuf<double> qs{size_t(3 * Q)}, rec2{size_t(Q)}, qavg;
        DevBuf<double4> rec{size_t(Q)};
        const bool nonlocal = any_nonlocal(gp);
        if (nonlocal) qavg.alloc(size_t(3 * Q));
        const int poly = influence_class(gp);
        if (m->nn == NLF_ELEMENT_Q4) {
            hipLaunchKernelGGL((k_thermal_stress<4, 4>), dim3(blocks_for(Q, 256)), dim3(256), 0, s, m->E, m->el.p, m->jac.p,
                               m->xq.p, m->group_of.p, d_groups.p, m->fe, delta_temperature_dev, qs.p, rec.p, rec2.p);
            if (nonlocal) launch_average<4, 3>(m, d_avg.p, poly, rec.p, rec2.p, qs.p, qavg.p, s);
            hipLaunchKernelGGL((k_thermal_load<4, 4>), dim3(blocks_for(m->N, 256)), dim3(256), 0, s, m->N, m->ne_off.p,
                               m->ne_ids.p, m->el.p, m->D.p, m->group_of.p, d_groups.p, m->fe, Q, qs.p,
                               nonlocal ? qavg.p : qs.p, f_dev);
        } else {
            hipLaunchKernelGGL((k_thermal_stress<8, 9>), dim3(blocks_for(Q, 256)), dim3(256), 0, s, m->E, m->el.p, m->jac.p,
                               m->xq.p, m->group_of.p, d_groups.p, m->fe, delta_temperature_dev, qs.p, rec.p, rec2.p);
            if (nonlocal) launch_average<9, 3>(m, d_avg.p, poly, rec.p, rec2.p, qs.p, qavg.p, s);
            hipLaunchKernelGGL((k_thermal_load<8, 9>), dim3(blocks_for(m->N, 256)), dim3(256), 0, s, m->N, m->ne_off.p,
                               m->ne_ids.p, m->el.p, m->D.p, m->group_of.p, d_groups.p, m->fe, Q, qs.p,
                               nonlocal ? qavg.p : qs.p, f_dev);
        }
        NLF_HIP(hipGetLastError());
        NLF_HIP(hipStreamSynchronize(s));
    });
}

// Heat flux of a temperature field: heat_equation_solution_2d::calc_flux
// (solvers/solver_2d/thermal/heat_equation_solution_2d.hpp:99-137).
//
//   1. k_flux_local (one thread per quadrature point): the gradient sum_i T_i D_i(q) / |J_q|
//      (gradient_in_qnodes, mesh_2d_utils.hpp:39-63) times -lambda of the element's group
//      (local_flux_in_qnodes, heat_equation_solution_2d.hpp:72-97).  Besides the local flux it
//      writes, per quadrature point, the record {x_q, y_q, w_q |J_q| q_x, w_q |J_q| q_y} that the
//      nonlocal gather streams.
//   2. k_flux_nonlocal (one wave per element eL): for a nonlocal group
//        q(x_L) = p1 q_loc(x_L) + (1 - p1) sum_{eNL in N(eL)} sum_qNL w |J| phi(x_L, x_NL) q_loc(x_NL)
//      (:106-131).  The lanes run over the (neighbour, quadrature point) pairs of eL, so
//      consecutive lanes read consecutive 32-byte records, and each lane evaluates phi against the
//      nq points of eL; a wave reduction gives the nq sums.  Local groups copy q_loc.
//   3. k_flux_nodes (one thread per node): qnodes_to_nodes (mesh_2d_utils.hpp:82-101), the
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
                                                    const double* __restrict__ T, double2* __restrict__ qloc,
                                                    double4* __restrict__ rec) {
    const i64 idx = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= E * NQ) return;
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
    qloc[idx] = make_double2(fx, fy);
    const double wj = t.w[q] * j;
    rec[idx] = make_double4(xq[2 * idx], xq[2 * idx + 1], wj * fx, wj * fy);
}

// phi for the flux gather.  POLY = 1 or 2: polynomial influence with p = n (d^(p/n) = d) and q = POLY
// (the fast polynomial, the bell), evaluated with the same operations as influence() -- the same
// division form near the cutoff, (1 - d) and its square as one multiply -- without the code of
// the other kinds, which keeps the gather loop small; POLY = 0: influence() itself.
template<int POLY>
__device__ __forceinline__ double flux_phi(const InfluenceParams& f, double ax, double ay, double bx, double by) {
    if (POLY == 0) return influence(f, ax, ay, bx, by);
    const double dx = ax - bx, dy = ay - by;
    const double zx = dx * f.inv_r[0], zy = dy * f.inv_r[1];
    double d = zx * zx + zy * zy;
    if (fabs(d - 1.0) < 1e-12) {
        const double ex = dx / f.r[0], ey = dy / f.r[1];
        d = ex * ex + ey * ey;
    }
    const double s = 1.0 - d;
    return d < 1.0 ? f.norm * (POLY == 1 ? s : s * s) : 0.0;
}

template<int NQ, int POLY>
__global__ __launch_bounds__(256) void k_flux_nonlocal(i64 E, i64 Q, const i64* __restrict__ nb_off,
                                                       const i32* __restrict__ nb_ids, const i32* __restrict__ group_of,
                                                       const GroupParams* __restrict__ groups,
                                                       const double4* __restrict__ rec,
                                                       const double2* __restrict__ qloc, double* __restrict__ qout) {
    // eL is the same for the whole wave; saying so lets the group parameters, the neighbour range
    // and every branch on the influence kind live in scalar registers
    const i64 eL = __builtin_amdgcn_readfirstlane(int(blockIdx.x) * 4 + int(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    if (eL >= E) return;
    const GroupParams& g = groups[__builtin_amdgcn_readfirstlane(group_of[eL])];
    if (!g.nonlocal) {
        if (lane < NQ) {
            const double2 f = qloc[eL * NQ + lane];
            qout[eL * NQ + lane] = f.x;
            qout[Q + eL * NQ + lane] = f.y;
        }
        return;
    }
    const InfluenceParams inf = g.inf;
    double xl[NQ], yl[NQ], ax[NQ], ay[NQ];
    for (int q = 0; q < NQ; ++q) {
        const double4 r = rec[eL * NQ + q];
        xl[q] = r.x;
        yl[q] = r.y;
        ax[q] = 0.0;
        ay[q] = 0.0;
    }
    const i64 b = nb_off[eL];
    const i64 pairs = (nb_off[eL + 1] - b) * NQ;
    // pair t = (neighbour t / NQ, its quadrature point t % NQ).  Software pipeline: while pair t is
    // evaluated, the record of pair t + 64 and the neighbour id of pair t + 128 are in flight, so
    // the id -> record load chain is not waited for in the loop.
    i64 t = lane;
    double4 r = make_double4(0.0, 0.0, 0.0, 0.0);
    i32 id_next = 0;
    if (t < pairs) r = rec[i64(nb_ids[b + t / NQ]) * NQ + t % NQ];
    if (t + 64 < pairs) id_next = nb_ids[b + (t + 64) / NQ];
    for (; t < pairs; t += 64) {
        double4 r_next = make_double4(0.0, 0.0, 0.0, 0.0);
        i32 id_after = 0;
        if (t + 64 < pairs) r_next = rec[i64(id_next) * NQ + (t + 64) % NQ];
        if (t + 128 < pairs) id_after = nb_ids[b + (t + 128) / NQ];
        for (int q = 0; q < NQ; ++q) {
            const double phi = flux_phi<POLY>(inf, xl[q], yl[q], r.x, r.y);
            ax[q] += phi * r.z;
            ay[q] += phi * r.w;
        }
        r = r_next;
        id_next = id_after;
    }
    for (int q = 0; q < NQ; ++q)
        for (int o = 32; o > 0; o >>= 1) {
            ax[q] += __shfl_xor(ax[q], o);
            ay[q] += __shfl_xor(ay[q], o);
        }
    if (lane < NQ) {
        double sx = 0.0, sy = 0.0;
        for (int q = 0; q < NQ; ++q)
            if (q == lane) { sx = ax[q]; sy = ay[q]; }
        const double2 f = qloc[eL * NQ + lane];
        const double p1 = g.p1, w1 = 1.0 - g.p1;
        qout[eL * NQ + lane] = f.x * p1 + sx * w1;
        qout[Q + eL * NQ + lane] = f.y * p1 + sy * w1;
    }
}

template<int NN, int NQ>
__global__ __launch_bounds__(256) void k_flux_nodes(i64 N, i64 Q, const i32* __restrict__ ne_off,
                                                    const i32* __restrict__ ne_ids, const i32* __restrict__ el,
                                                    const double* __restrict__ jac, FeTables t,
                                                    const double* __restrict__ qout, double* __restrict__ out) {
    const i64 node = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (node >= N) return;
    double vx = 0.0, vy = 0.0, node_area = 0.0;
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
        vx += area * qout[qs];
        vy += area * qout[Q + qs];
        node_area += area;
    }
    out[node] = vx / node_area;
    out[N + node] = vy / node_area;
}

template<int NN, int NQ>
void launch_flux(nlf_mesh m, const GroupParams* groups, int poly, const double* T, double2* qloc, double4* rec,
                 double* qout, double* out, hipStream_t s) {
    const i64 Q = m->E * NQ;
    hipLaunchKernelGGL((k_flux_local<NN, NQ>), dim3(blocks_for(Q, 256)), dim3(256), 0, s, m->E, m->el.p, m->D.p,
                       m->jac.p, m->xq.p, m->group_of.p, groups, m->fe, T, qloc, rec);
    NLF_HIP(hipGetLastError());
    auto nonlocal = poly == 2 ? k_flux_nonlocal<NQ, 2> : (poly == 1 ? k_flux_nonlocal<NQ, 1> : k_flux_nonlocal<NQ, 0>);
    hipLaunchKernelGGL(nonlocal, dim3(blocks_for(m->E, 4)), dim3(256), 0, s, m->E, Q,
                       m->has_neighbours ? m->nb_off.p : nullptr, m->has_neighbours ? m->nb_ids.p : nullptr,
                       m->group_of.p, groups, rec, qloc, qout);
    NLF_HIP(hipGetLastError());
    if (out) {
        hipLaunchKernelGGL((k_flux_nodes<NN, NQ>), dim3(blocks_for(m->N, 256)), dim3(256), 0, s, m->N, Q, m->ne_off.p,
                           m->ne_ids.p, m->el.p, m->jac.p, m->fe, qout, out);
        NLF_HIP(hipGetLastError());
    }
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
        bool any_nonlocal = false;
        int poly = -1;                       // influence class of the nonlocal groups (flux_phi)
        for (int g = 0; g < m->G; ++g) {
            gp[g] = model_group_params(models[g], false);
            const nlf_thermal_material& t = materials[g];
            NLF_CHECK(t.kind >= NLF_CONDUCTIVITY_ISOTROPIC && t.kind <= NLF_CONDUCTIVITY_ANISOTROPIC, NLF_ERR_ARGUMENT,
                      "unknown conductivity kind");
            gp[g].ckind = t.kind;
            gp[g].c[0] = t.c[0]; gp[g].c[1] = t.c[1]; gp[g].c[2] = t.c[2];
            any_nonlocal |= gp[g].nonlocal != 0;
            // the specialised gather needs one polynomial class for all nonlocal groups
            const InfluenceParams& f = gp[g].inf;
            if (gp[g].nonlocal) {
                int cls = 0;
                if (f.kind == NLF_INFLUENCE_POLYNOMIAL && (f.fast || (f.p_int && f.p_i == 1 && f.q_int && f.q_i <= 2)))
                    cls = f.fast ? 1 : f.q_i;
                poly = poly < 0 ? cls : (poly == cls ? poly : 0);
            }
        }
        NLF_CHECK(!any_nonlocal || m->has_neighbours, NLF_ERR_STATE,
                  "the flux of a nonlocal group needs the neighbour lists (nlf_find_neighbours)");
        if (m->E == 0) return;
        const i64 Q = m->E * m->fe.nq;
        DevBuf<GroupParams> d_groups;
        upload(d_groups, gp.data(), gp.size(), s);
        DevBuf<double2> qloc{size_t(Q)};
        DevBuf<double4> rec{size_t(Q)};
        DevBuf<double> own;
        double* qout = qflux_dev;
        if (!qout) {
            own.alloc(size_t(2 * Q));
            qout = own.p;
        }
        if (m->nn == NLF_ELEMENT_Q4) launch_flux<4, 4>(m, d_groups.p, poly, T_dev, qloc.p, rec.p, qout, flux_dev, s);
        else launch_flux<8, 9>(m, d_groups.p, poly, T_dev, qloc.p, rec.p, qout, flux_dev, s);
        NLF_HIP(hipStreamSynchronize(s));
    });
}

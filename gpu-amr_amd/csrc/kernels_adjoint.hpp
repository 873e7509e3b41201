// kernels_adjoint.hpp -- discrete adjoint of the advection step, the halo
// fill and the regrid transfer (their exact transposes), in gather form so
// that every sum has a fixed order and no atomics are needed.
//
// For fixed topology, velocity v and dt, one step is linear:
//   u' = u - sum_d dt/dx_d [ F(u, u_R) - F(u_L, u) ],
//   F(a, b) = 0.5 (a v + b v) - 0.5 |v| (b - a) = vp a + vm b,
//   vp = 0.5 (v + |v|), vm = 0.5 (v - |v|),
// so  du'/du = 1 - sum_d dt/dx_d |v_d|,  du'/du_L = dt/dx vp,
//     du'/du_R = -dt/dx vm,
// and du'/dv_d = -dt/dx_d [ G(u, u_R) - G(u_L, u) ],
//     G(a, b) = 0.5 (a + b) - 0.5 sgn(v_d) (b - a)
// (dt held fixed).  Ghosts are linear in the neighbours' interiors:
// same = copy, coarser = injection, finer = mean of 2^R fine cells.  Their
// transposes: a cell copied into a same-level neighbour's ghost receives that
// ghost's adjoint; a cell injected into finer neighbours' ghosts receives the
// sum of their adjoints; a cell averaged into a coarser neighbour's ghost
// receives 1/2^R of it.  The adjoint of a ghost is the coefficient with
// which the neighbour's boundary cell reads it, times that cell's lambda'.
//
// One adjoint step is one launch over the leaves of the step's topology
// (k_adj_step): lambda = A^T lambda' with the transposed halo gathered from
// the neighbours' boundary layers of lambda', plus the per-leaf partial sums
// of lambda' . du'/dv.  A regrid transfer is transposed by k_adj_regrid.
#pragma once

#include "device_common.hpp"
#include "kernels_step.hpp"

#pragma clang fp contract(off)

namespace amrhip
{

// Per-launch adjoint constants: dt/dx per level and dimension, vp, vm and
// sgn(v) per dimension.
struct adjoint_params
{
    double dtdx[kMaxLevels][3];
    double vp[3];
    double vm[3];
    double sgn[3];
};

// In-patch part of (A^T lambda')_c.
template <int R>
__host__ __device__ inline double adj_interior(box<R> const& B, const double* __restrict__ lam, int i,
                                               const double* dtdx, const double* vp, const double* vm)
{
    int c[R];
    B.coords(i, c);
    double diag = 1.0;
    for (int d = 0; d < R; ++d) diag -= dtdx[d] * (vp[d] - vm[d]);
    double acc = diag * lam[i];
    for (int d = 0; d < R; ++d)
    {
        const int j = R - 1 - d; // layout dim of physical axis d
        const int s = B.stride(j);
        if (c[j] + 1 < B.n(j)) acc += (dtdx[d] * vp[d]) * lam[i + s]; // c is u_L of c + e_d
        if (c[j] - 1 >= 0) acc -= (dtdx[d] * vm[d]) * lam[i - s];     // c is u_R of c - e_d
    }
    return acc;
}

template <int R, int N>
__host__ __device__ inline double adj_interior(const double* __restrict__ lam, int i, const double* dtdx,
                                               const double* vp, const double* vm)
{
    constexpr box<R> B = box<R>::cube(N);
    return adj_interior<R>(B, lam, i, dtdx, vp, vm);
}

// Transposed halo term for the leaf's boundary cells facing direction dir at
// tangential layout coordinates c (c[dim] ignored), from the neighbours'
// lambda'.  It is added to the depth-0 cell, and for a coarser neighbour also
// to the depth-1 cell (both are averaged into the same coarse ghost).
//   same    : the neighbour reads our cell as its ghost on side dir^1;
//   finer   : the 2^(R-1) fine ghost cells injected from our cell;
//   coarser : the coarse ghost cell that averages our cell, / 2^R.
// The neighbour's boundary cell reads its ghost on side dir^1 with
// coefficient dt/dx vp (ghost is its u_L, dir positive) or -dt/dx vm (ghost
// is its u_R, dir negative), at the neighbour's own level.
template <int R>
__host__ __device__ inline double adj_face_term(box<R> const& B, const double* __restrict__ lam, const face_link& L,
                                                uint32_t meta, int dir, const int* c, int level,
                                                const adjoint_params& prm)
{
    const int  dim = dir >> 1;
    const bool pos = (dir & 1) != 0;
    const int  d   = R - 1 - dim;
    const int  nb  = pos ? 0 : B.n(dim) - 1; // the neighbour's layer that touches us
    switch (rel_of(meta))
    {
        case REL_SAME:
        {
            int cc[R];
            for (int k = 0; k < R; ++k) cc[k] = (k == dim) ? nb : c[k];
            const double coef = pos ? (prm.dtdx[level][d] * prm.vp[d]) : -(prm.dtdx[level][d] * prm.vm[d]);
            return coef * lam[static_cast<size_t>(L.ids[0]) * B.ni() + B.lin(cc)];
        }
        case REL_FINER:
        {
            int kid = 0, bit = 0, base[R];
            for (int k = 0; k < R; ++k)
            {
                if (k == dim)
                {
                    base[k] = nb;
                    continue;
                }
                kid |= (c[k] / B.half(k)) << bit;
                ++bit;
                base[k] = (c[k] % B.half(k)) * 2;
            }
            const double* q   = lam + static_cast<size_t>(L.ids[kid]) * B.ni();
            double        sum = 0.0;
            for (int o = 0; o < box<R>::K; ++o)
            {
                int cc[R], b = 0;
                for (int k = 0; k < R; ++k)
                {
                    if (k == dim)
                    {
                        cc[k] = base[k];
                        continue;
                    }
                    cc[k] = base[k] + ((o >> (R - 2 - b)) & 1);
                    ++b;
                }
                sum += q[B.lin(cc)];
            }
            const double coef =
                pos ? (prm.dtdx[level + 1][d] * prm.vp[d]) : -(prm.dtdx[level + 1][d] * prm.vm[d]);
            return coef * sum;
        }
        case REL_COARSER:
        {
            int cc[R];
            for (int k = 0; k < R; ++k)
                cc[k] = (k == dim) ? nb : static_cast<int>(quad_of(meta, k)) * B.half(k) + c[k] / 2;
            const double coef =
                pos ? (prm.dtdx[level - 1][d] * prm.vp[d]) : -(prm.dtdx[level - 1][d] * prm.vm[d]);
            return coef * lam[static_cast<size_t>(L.ids[0]) * B.ni() + B.lin(cc)] / static_cast<double>(box<R>::F);
        }
        default: return 0.0;
    }
}

template <int R, int N>
__host__ __device__ inline double adj_face_term(const double* __restrict__ lam, const face_link& L, uint32_t meta,
                                                int dir, const int* c, int level, const adjoint_params& prm)
{
    constexpr box<R> B = box<R>::cube(N);
    return adj_face_term<R>(B, lam, L, meta, dir, c, level, prm);
}

// G-form contribution of a cell to lambda' . du'/dv along physical axis d,
// from its value and its two neighbours.
__host__ __device__ inline double adj_velocity_term(double uc, double ul, double ur, double lam_c, double dtdx,
                                                    double sgn)
{
    const double gR = 0.5 * (uc + ur) - 0.5 * sgn * (ur - uc);
    const double gL = 0.5 * (ul + uc) - 0.5 * sgn * (uc - ul);
    return lam_c * (-dtdx * (gR - gL));
}

// Contribution of cell i to lambda' . du'/dv, from the staged u tile (with
// the ghost layer).
template <int R, int N>
__host__ __device__ inline void adj_velocity_cell(const double* tile, int i, double lam_c, const double* dtdx,
                                                  const double* sgn, double* acc)
{
    using T = tile_shape<R, N>;
    int c[R];
    patch_shape<R, N>::coords(i, c);
    const int    t  = T::idx(c);
    const double uc = tile[t];
    for (int d = 0; d < R; ++d)
    {
        const int s = T::stride(R - 1 - d);
        acc[d] += adj_velocity_term(uc, tile[t - s], tile[t + s], lam_c, dtdx[d], sgn[d]);
    }
}

// adj_velocity_cell with the neighbours read from memory (any patch shape).
template <int R>
__host__ __device__ inline void adj_velocity_cell_direct(box<R> const& B, const double* __restrict__ u, uint32_t leaf,
                                                         const face_link* __restrict__ links,
                                                         const uint32_t* __restrict__ metas, int i, double lam_c,
                                                         const double* dtdx, const double* sgn, double* acc)
{
    int c[R];
    B.coords(i, c);
    const double uc = u[static_cast<size_t>(leaf) * B.ni() + i];
    for (int d = 0; d < R; ++d)
    {
        const int j = R - 1 - d;
        c[j] -= 1;
        const double ul = cell_or_ghost<R>(B, u, leaf, links, metas, c);
        c[j] += 2;
        const double ur = cell_or_ghost<R>(B, u, leaf, links, metas, c);
        c[j] -= 1;
        acc[d] += adj_velocity_term(uc, ul, ur, lam_c, dtdx[d], sgn[d]);
    }
}

// Face index of tangential coordinates c (dims other than dim), row-major.
template <int R>
__host__ __device__ inline int face_index(box<R> const& B, const int* c, int dim)
{
    int f = 0;
    for (int k = 0; k < R; ++k)
        if (k != dim) f = f * B.n(k) + c[k];
    return f;
}

template <int R, int N>
__host__ __device__ inline int face_index(const int* c, int dim)
{
    constexpr box<R> B = box<R>::cube(N);
    return face_index<R>(B, c, dim);
}

// Tangential coordinates of face index f of direction dir (c[dim] = 0).
template <int R>
__host__ __device__ inline void face_coords(box<R> const& B, int f, int dir, int* c)
{
    const int dim = dir >> 1;
    for (int k = R - 1; k >= 0; --k)
    {
        if (k == dim)
        {
            c[k] = 0;
            continue;
        }
        c[k] = f % B.n(k);
        f /= B.n(k);
    }
}

template <int R, int N>
__host__ __device__ inline void face_coords(int f, int dir, int* c)
{
    constexpr box<R> B = box<R>::cube(N);
    face_coords<R>(B, f, dir, c);
}

// lambda_c = (A^T lambda')_c + transposed halo terms h[dir][face] of the
// directions whose boundary band holds c.
template <int R, int N>
__host__ __device__ inline double adj_cell(const double* __restrict__ lam, const double* __restrict__ h,
                                           const uint32_t* __restrict__ M, int i, const double* dtdx,
                                           const double* vp, const double* vm)
{
    using S    = patch_shape<R, N>;
    double acc = adj_interior<R, N>(lam, i, dtdx, vp, vm);
    int    c[R];
    S::coords(i, c);
    for (int dir = 0; dir < S::NDIR; ++dir)
    {
        const int dim   = dir >> 1;
        const int depth = (dir & 1) ? (N - 1 - c[dim]) : c[dim];
        if (depth == 0 || (depth == 1 && rel_of(M[dir]) == REL_COARSER))
            acc += h[dir * S::FACE + face_index<R, N>(c, dim)];
    }
    return acc;
}

// adj_cell with lambda' and the transposed halo terms read from memory (any
// patch shape; the terms are added in adj_cell's order, so the same bits).
template <int R>
__host__ __device__ inline double adj_cell_direct(box<R> const& B, const double* __restrict__ lam_next, uint32_t leaf,
                                                  const face_link* __restrict__ L, const uint32_t* __restrict__ M,
                                                  int i, int level, const adjoint_params& prm, const double* dtdx,
                                                  const double* vp, const double* vm)
{
    double acc = adj_interior<R>(B, lam_next + static_cast<size_t>(leaf) * B.ni(), i, dtdx, vp, vm);
    int    c[R];
    B.coords(i, c);
    for (int dir = 0; dir < box<R>::NDIR; ++dir)
    {
        const int dim   = dir >> 1;
        const int depth = (dir & 1) ? (B.n(dim) - 1 - c[dim]) : c[dim];
        if (depth == 0 || (depth == 1 && rel_of(M[dir]) == REL_COARSER))
            acc += adj_face_term<R>(B, lam_next, L[dir], M[dir], dir, c, level, prm);
    }
    return acc;
}

// Transposed regrid transfer, value of cell t of old leaf i.
// Refined leaf: sum of lambda over the 2^R fine cells injected from t.
template <int R>
__host__ __device__ inline double adj_refine_value(box<R> const& B, const double* __restrict__ lam_new,
                                                   uint32_t first_child, int t)
{
    int tc[R], base[R], kid = 0;
    B.coords(t, tc);
    for (int d = 0; d < R; ++d)
    {
        kid |= (tc[d] / B.half(d)) << (R - 1 - d);
        base[d] = (tc[d] % B.half(d)) * 2;
    }
    const double* q   = lam_new + static_cast<size_t>(first_child + kid) * B.ni();
    double        sum = 0.0;
    for (int off = 0; off < box<R>::F; ++off)
    {
        int cc[R];
        for (int d = 0; d < R; ++d) cc[d] = base[d] + ((off >> (R - 1 - d)) & 1);
        sum += q[B.lin(cc)];
    }
    return sum;
}

// Coarsened family: cell f of child kid receives lambda of the parent cell
// that averaged it, divided by 2^R.
template <int R>
__host__ __device__ inline double adj_restrict_value(box<R> const& B, const double* __restrict__ lam_parent, int kid,
                                                     int f)
{
    int fc[R], t[R];
    B.coords(f, fc);
    for (int d = 0; d < R; ++d) t[d] = ((kid >> (R - 1 - d)) & 1) * B.half(d) + fc[d] / 2;
    return lam_parent[B.lin(t)] / static_cast<double>(box<R>::F);
}

template <int R, int N>
__host__ __device__ inline double adj_refine_value(const double* __restrict__ lam_new, uint32_t first_child, int t)
{
    constexpr box<R> B = box<R>::cube(N);
    return adj_refine_value<R>(B, lam_new, first_child, t);
}

template <int R, int N>
__host__ __device__ inline double adj_restrict_value(const double* __restrict__ lam_parent, int kid, int f)
{
    constexpr box<R> B = box<R>::cube(N);
    return adj_restrict_value<R>(B, lam_parent, kid, f);
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// One adjoint step, one workgroup per leaf: lam = transpose(step)(lam_next)
// and vpart[leaf][d] = sum over the leaf's cells of lam_next . du'/dv_d.  u
// is the state the forward step started from; the step's constants are read
// from prm_p (device memory, so a captured launch picks up new values).
template <int R, int N, int NT>
__global__ __launch_bounds__(NT) void k_adj_step(const double* __restrict__ lam_next, const double* __restrict__ u,
                                                 const uint64_t* __restrict__ keys,
                                                 const face_link* __restrict__ links,
                                                 const uint32_t* __restrict__ metas,
                                                 const adjoint_params* __restrict__ prm_p, double* __restrict__ lam,
                                                 double* __restrict__ vpart, uint32_t nleaves)
{
    const adjoint_params& prm = *prm_p;
    using S = patch_shape<R, N>;
    __shared__ double tile[tile_shape<R, N>::TS];
    __shared__ double lt[S::NI];
    __shared__ double h[S::NDIR * S::FACE];
    constexpr int     NW = NT / 64;
    static_assert(NT % 64 == 0, "whole waves");
    __shared__ double red[R][NW];
    const uint32_t leaf = blockIdx.x;
    if (leaf >= nleaves) return;
    const face_link* L  = links + static_cast<size_t>(leaf) * S::NDIR;
    const uint32_t*  M  = metas + static_cast<size_t>(leaf) * S::NDIR;
    const int        lv = key_level(keys[leaf]);
    stage_tile<R, N, NT>(tile, u, leaf, links, metas);
    const double* lp = lam_next + static_cast<size_t>(leaf) * S::NI;
    for (int i = threadIdx.x; i < S::NI; i += NT) lt[i] = lp[i];
    // The face terms start half a workgroup away from the ghost gather of
    // stage_tile (which starts at thread 0), so that their two dependent
    // load chains run on different waves.
    for (int g = (threadIdx.x + NT / 2) % NT; g < S::NDIR * S::FACE; g += NT)
    {
        const int dir = g / S::FACE;
        int       c[R];
        face_coords<R, N>(g % S::FACE, dir, c);
        h[g] = adj_face_term<R, N>(lam_next, L[dir], M[dir], dir, c, lv, prm);
    }
    __syncthreads();
    double dtdx[R], vp[R], vm[R], sgn[R], acc[R];
    for (int d = 0; d < R; ++d)
    {
        dtdx[d] = prm.dtdx[lv][d];
        vp[d]   = prm.vp[d];
        vm[d]   = prm.vm[d];
        sgn[d]  = prm.sgn[d];
        acc[d]  = 0.0;
    }
    double* out = lam + static_cast<size_t>(leaf) * S::NI;
    for (int i = threadIdx.x; i < S::NI; i += NT)
    {
        out[i] = adj_cell<R, N>(lt, h, M, i, dtdx, vp, vm);
        adj_velocity_cell<R, N>(tile, i, lt[i], dtdx, sgn, acc);
    }
    // fixed-order reduction of the velocity partials: a shuffle tree inside
    // each wave, then the waves' sums in wave order
    for (int d = 0; d < R; ++d)
        for (int o = 32; o > 0; o >>= 1) acc[d] += __shfl_down(acc[d], o, 64);
    if ((threadIdx.x & 63) == 0)
        for (int d = 0; d < R; ++d) red[d][threadIdx.x >> 6] = acc[d];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int d = 0; d < 3; ++d)
        {
            double sum = 0.0;
            if (d < R)
            {
                sum = red[d][0];
                for (int w = 1; w < NW; ++w) sum += red[d][w];
            }
            vpart[static_cast<size_t>(leaf) * 3 + d] = sum;
        }
}

// Transposed regrid transfer of old leaf i (one workgroup; mark / head /
// counts / pos recorded by the forward regrid).
template <int R, int NT>
__device__ inline void adj_regrid_leaf(box<R> const& B, const double* __restrict__ lam_new,
                                       const uint8_t* __restrict__ mark, const uint8_t* __restrict__ head,
                                       const uint32_t* __restrict__ counts, const uint32_t* __restrict__ pos,
                                       double* __restrict__ lam_old, uint32_t i)
{
    if (counts[i] == 0) return;
    const int      ni  = B.ni();
    const uint32_t dst = pos[i];
    double*        out = lam_old + static_cast<size_t>(i) * ni;
    if (mark[i])
    {
        for (int t = threadIdx.x; t < ni; t += NT) out[t] = adj_refine_value<R>(B, lam_new, dst, t);
    }
    else if (head[i])
    {
        const double* par = lam_new + static_cast<size_t>(dst) * ni;
        for (int kid = 0; kid < box<R>::F; ++kid)
            for (int f = threadIdx.x; f < ni; f += NT)
                lam_old[static_cast<size_t>(i + kid) * ni + f] = adj_restrict_value<R>(B, par, kid, f);
    }
    else
    {
        const double* src = lam_new + static_cast<size_t>(dst) * ni;
        for (int t = threadIdx.x; t < ni; t += NT) out[t] = src[t];
    }
}

// Transposed regrid transfer, one workgroup per leaf before the regrid.
template <int R, int N, int NT>
__global__ __launch_bounds__(NT) void k_adj_regrid(const double* __restrict__ lam_new, const uint8_t* __restrict__ mark,
                                                   const uint8_t* __restrict__ head,
                                                   const uint32_t* __restrict__ counts,
                                                   const uint32_t* __restrict__ pos, double* __restrict__ lam_old,
                                                   uint32_t n_old)
{
    constexpr box<R> B = box<R>::cube(N);
    const uint32_t   i = blockIdx.x;
    if (i >= n_old) return;
    adj_regrid_leaf<R, NT>(B, lam_new, mark, head, counts, pos, lam_old, i);
}

// k_adj_regrid for patches of any shape.
template <int R, int NT>
__global__ __launch_bounds__(NT) void k_adj_regrid_box(box<R> B, const double* __restrict__ lam_new,
                                                       const uint8_t* __restrict__ mark,
                                                       const uint8_t* __restrict__ head,
                                                       const uint32_t* __restrict__ counts,
                                                       const uint32_t* __restrict__ pos, double* __restrict__ lam_old,
                                                       uint32_t n_old)
{
    const uint32_t i = blockIdx.x;
    if (i >= n_old) return;
    adj_regrid_leaf<R, NT>(B, lam_new, mark, head, counts, pos, lam_old, i);
}

// k_adj_step for patches of any shape: lambda', u and the transposed halo
// terms read from memory (the operands and sums of k_adj_step).
template <int R, int NT>
__global__ __launch_bounds__(NT) void k_adj_step_box(box<R> B, const double* __restrict__ lam_next,
                                                     const double* __restrict__ u, const uint64_t* __restrict__ keys,
                                                     const face_link* __restrict__ links,
                                                     const uint32_t* __restrict__ metas,
                                                     const adjoint_params* __restrict__ prm_p,
                                                     double* __restrict__ lam, double* __restrict__ vpart,
                                                     uint32_t nleaves)
{
    const adjoint_params& prm = *prm_p;
    constexpr int NW = NT / 64;
    static_assert(NT % 64 == 0, "whole waves");
    __shared__ double red[R][NW];
    const uint32_t    leaf = blockIdx.x;
    if (leaf >= nleaves) return;
    const face_link* L  = links + static_cast<size_t>(leaf) * box<R>::NDIR;
    const uint32_t*  M  = metas + static_cast<size_t>(leaf) * box<R>::NDIR;
    const int        lv = key_level(keys[leaf]);
    double           dtdx[R], vp[R], vm[R], sgn[R], acc[R];
    for (int d = 0; d < R; ++d)
    {
        dtdx[d] = prm.dtdx[lv][d];
        vp[d]   = prm.vp[d];
        vm[d]   = prm.vm[d];
        sgn[d]  = prm.sgn[d];
        acc[d]  = 0.0;
    }
    const int     ni  = B.ni();
    const double* lp  = lam_next + static_cast<size_t>(leaf) * ni;
    double*       out = lam + static_cast<size_t>(leaf) * ni;
    for (int i = threadIdx.x; i < ni; i += NT)
    {
        out[i] = adj_cell_direct<R>(B, lam_next, leaf, L, M, i, lv, prm, dtdx, vp, vm);
        adj_velocity_cell_direct<R>(B, u, leaf, links, metas, i, lp[i], dtdx, sgn, acc);
    }
    // the velocity partials reduced as in k_adj_step
    for (int d = 0; d < R; ++d)
        for (int o = 32; o > 0; o >>= 1) acc[d] += __shfl_down(acc[d], o, 64);
    if ((threadIdx.x & 63) == 0)
        for (int d = 0; d < R; ++d) red[d][threadIdx.x >> 6] = acc[d];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int d = 0; d < 3; ++d)
        {
            double sum = 0.0;
            if (d < R)
            {
                sum = red[d][0];
                for (int w = 1; w < NW; ++w) sum += red[d][w];
            }
            vpart[static_cast<size_t>(leaf) * 3 + d] = sum;
        }
}

// The velocity partials of a run of adjoint steps, one workgroup per step:
// step first + b wrote its leaves' triples at vpart + b * stride * 3 (leaves[]
// per step); they are summed in one fixed order (leaves t, t + 256, ... per
// thread, a shuffle tree inside each wave, the waves in order) into that
// step's slot of step_sums.
__global__ __launch_bounds__(256) void k_sum3_steps(const double* __restrict__ vpart, uint64_t stride,
                                                    const uint32_t* __restrict__ leaves, uint64_t first,
                                                    double* __restrict__ step_sums)
{
    __shared__ double red[3][4];
    const uint64_t    step = first + blockIdx.x;
    const uint32_t    n    = leaves[step];
    const double*     v    = vpart + blockIdx.x * stride * 3;
    double            s[3] = { 0.0, 0.0, 0.0 };
    for (uint32_t l = threadIdx.x; l < n; l += 256)
        for (int d = 0; d < 3; ++d) s[d] += v[static_cast<size_t>(l) * 3 + d];
    for (int d = 0; d < 3; ++d)
        for (int o = 32; o > 0; o >>= 1) s[d] += __shfl_down(s[d], o, 64);
    if ((threadIdx.x & 63) == 0)
        for (int d = 0; d < 3; ++d) red[d][threadIdx.x >> 6] = s[d];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int d = 0; d < 3; ++d) step_sums[step * 3 + d] = ((red[d][0] + red[d][1]) + red[d][2]) + red[d][3];
}

// Deterministic sum of `count` triples: stage 1 gives each of kReduceBlocks
// workgroups a fixed contiguous chunk, stage 2 adds the chunk sums in order
// into acc.
constexpr int kReduceBlocks = 64;

__device__ inline void block_sum3(double* s, double (*red)[256])
{
    for (int d = 0; d < 3; ++d) red[d][threadIdx.x] = s[d];
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1)
    {
        if (static_cast<int>(threadIdx.x) < k)
            for (int d = 0; d < 3; ++d) red[d][threadIdx.x] += red[d][threadIdx.x + k];
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_sum3_chunks(const double* __restrict__ v, uint64_t count,
                                                     double* __restrict__ chunk_sums)
{
    __shared__ double red[3][256];
    const uint64_t    per = (count + kReduceBlocks - 1) / kReduceBlocks;
    const uint64_t    b0  = blockIdx.x * per;
    const uint64_t    b1  = b0 + per < count ? b0 + per : count;
    double            s[3] = { 0.0, 0.0, 0.0 };
    for (uint64_t l = b0 + threadIdx.x; l < b1; l += 256)
        for (int d = 0; d < 3; ++d) s[d] += v[l * 3 + d];
    block_sum3(s, red);
    if (threadIdx.x == 0)
        for (int d = 0; d < 3; ++d) chunk_sums[blockIdx.x * 3 + d] = red[d][0];
}

__global__ __launch_bounds__(256) void k_sum3_final(const double* __restrict__ chunk_sums, double* __restrict__ acc)
{
    __shared__ double red[3][256];
    double            s[3] = { 0.0, 0.0, 0.0 };
    if (threadIdx.x < kReduceBlocks)
        for (int d = 0; d < 3; ++d) s[d] = chunk_sums[threadIdx.x * 3 + d];
    block_sum3(s, red);
    if (threadIdx.x == 0)
        for (int d = 0; d < 3; ++d) acc[d] = red[d][0];
}

} // namespace amrhip

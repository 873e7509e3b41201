// kernels_step.hpp -- the leaf-patch upwind update with the halo gather fused
// in, the padded (halo-exchanged) view, and the refinement criteria.
//
// State in HBM: one buffer of n^R interior doubles per leaf slot (compact,
// 4 KiB per 8^3 patch, 2 KiB per 16^2 patch), leaves in Morton order.  Ghost
// values are not stored: each workgroup gathers the ghost layer of its patch
// from the neighbours' interiors through the device neighbour table, with the
// reference's same / finer (mean) / coarser (injection) rules, into LDS.
//
// Every per-item body is __host__ __device__; the __global__ kernels only
// distribute items over threads, so the same bodies can be run on the CPU.
#pragma once

#include "device_common.hpp"

#pragma clang fp contract(off)

namespace amrhip
{

// Per-launch step constants.  dtdx[l][d] = dt / cell_size(l)[d] and
// hs[d] = 0.5 * |v_d| are formed on the host with the same IEEE operations
// the reference performs per cell (amr_solver.hpp:333,
// AdvectionPhysics.hpp:61-65).
struct step_params
{
    double   vel[3];
    double   hs[3];
    double   dtdx[kMaxLevels][3];
    uint32_t reverse; // walk this launch's leaves backwards (sweep_item)
};

// Leaf (or run) of workgroup b in a launch over `count` items: XCD-contiguous
// ranges (xcd_item), walked backwards when `reverse` is set.  Consecutive
// steps alternate the direction, so that a step starts on the leaves the
// previous one wrote last -- the lines most likely still held by the
// memory-side cache.  Items >= count are skipped by the caller.
__device__ inline uint32_t sweep_item(uint32_t b, uint32_t nblocks, uint32_t count, uint32_t reverse)
{
    const uint32_t i = xcd_item(b, nblocks);
    return (reverse && i < count) ? count - 1 - i : i;
}

// Output store of a step kernel.  The non-temporal form streams the new
// state past L2 so that L2 keeps the input lines that neighbouring leaves'
// ghost gathers re-read.
template <bool NT>
__device__ inline void store_cell(double* p, double v)
{
    if constexpr (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

// Step constants of a leaf: the launch's own, or, for a batch of
// concatenated trees, those of the leaf's instance.
__device__ inline const step_params& params_for(const step_params& prm, const step_params* __restrict__ table,
                                                const uint16_t* __restrict__ inst, uint32_t leaf)
{
    return table ? table[inst[leaf]] : prm;
}

// Ghost value of the cell with layout coordinates c (c[dim] lies outside
// [0, N): >= N on the positive side, < 0 on the negative side) of the leaf
// whose direction-`dir` neighbour record is (L, meta).  Restates
// halo_exchange_impl_t::{same_t, finer_t, coarser_t}
// (include/ndtree/patch_utils.hpp:315-441) in compact coordinates:
//   same    -> neighbour cell at the wrapped normal coordinate;
//   finer   -> mean of the 2^R fine cells, summed from 0.0 in row-major order;
//   coarser -> injection of the coarse cell under the contact quadrant.
template <int R, int N>
__host__ __device__ inline double ghost_value(const double* __restrict__ u, const face_link& L, uint32_t meta, int dir,
                                              const int* c)
{
    using S        = patch_shape<R, N>;
    const int  dim = dir >> 1;
    const bool pos = (dir & 1) != 0;
    const int  nrm = pos ? c[dim] - N : c[dim] + N; // reference's from_idxs[dim]
    switch (rel_of(meta))
    {
        case REL_SAME:
        {
            int cc[R];
            for (int d = 0; d < R; ++d) cc[d] = (d == dim) ? nrm : c[d];
            return u[static_cast<size_t>(L.ids[0]) * S::NI + S::lin(cc)];
        }
        case REL_COARSER:
        {
            int cc[R];
            for (int d = 0; d < R; ++d)
                cc[d] = static_cast<int>(quad_of(meta, d)) * S::HALF + ((d == dim) ? nrm : c[d]) / 2;
            return u[static_cast<size_t>(L.ids[0]) * S::NI + S::lin(cc)];
        }
        case REL_FINER:
        {
            int k = 0, bit = 0, base[R];
            for (int d = 0; d < R; ++d)
            {
                const int v = (d == dim) ? nrm : c[d];
                base[d]     = (v * 2) % N;
                if (d != dim)
                {
                    k |= (c[d] / S::HALF) << bit;
                    ++bit;
                }
            }
            const double* q   = u + static_cast<size_t>(L.ids[k]) * S::NI;
            double        sum = 0.0;
            for (int o = 0; o < S::F; ++o)
            {
                int cc[R];
                for (int d = 0; d < R; ++d) cc[d] = base[d] + ((o >> (R - 1 - d)) & 1);
                sum += q[S::lin(cc)];
            }
            return sum / static_cast<double>(S::F);
        }
        default: return 0.0;
    }
}

// LDS tile of the patch plus one ghost layer: (N+2)^R doubles.
template <int R, int N>
struct tile_shape
{
    static constexpr int TP = N + 2;
    static constexpr int TS = (R == 2) ? TP * TP : TP * TP * TP;
    __host__ __device__ static int idx(const int* c)
    {
        int l = 0;
        for (int d = 0; d < R; ++d) l = l * TP + (c[d] + 1);
        return l;
    }
    // stride of layout dim j inside the tile
    __host__ __device__ static constexpr int stride(int j)
    {
        return (R == 3 && j == 0) ? TP * TP : (j == R - 1 ? 1 : TP);
    }
};

// Interior cell i of the leaf into the tile.
template <int R, int N>
__host__ __device__ inline void stage_interior(double* tile, const double* __restrict__ up, int i)
{
    int c[R];
    patch_shape<R, N>::coords(i, c);
    tile[tile_shape<R, N>::idx(c)] = up[i];
}

// Ghost item g in [0, 2R * N^(R-1)): direction g / N^(R-1), tangential
// position g % N^(R-1) in row-major order over the non-normal dims.
template <int R, int N>
__host__ __device__ inline void stage_ghost(double* tile, const double* __restrict__ u, uint32_t leaf,
                                            const face_link* __restrict__ links, const uint32_t* __restrict__ metas,
                                            int g)
{
    using S       = patch_shape<R, N>;
    const int dir = g / S::FACE, dim = dir >> 1;
    int       c[R], r = g % S::FACE;
    for (int d = R - 1; d >= 0; --d)
    {
        if (d == dim) continue;
        c[d] = r % N;
        r /= N;
    }
    c[dim]                          = (dir & 1) ? N : -1;
    const size_t e                  = static_cast<size_t>(leaf) * S::NDIR + dir;
    tile[tile_shape<R, N>::idx(c)] = ghost_value<R, N>(u, links[e], metas[e], dir, c);
}

// x-face side buffers: per leaf slot, the interior layers x = 0 (side 0) and
// x = N-1 (side 1), N^(R-1) doubles each, row-major over the other layout
// dims.  An x face of a row-major patch touches every row of it (every
// 64-128 B line); the side buffer holds the same values contiguously, so a
// kernel that gathers its x ghosts from there reads 2 x N^(R-1) doubles
// instead of touching two whole neighbour patches.  A step kernel that reads
// them writes those of the state it produces.
template <int R, int N>
__host__ __device__ inline int xface_index(const int* c)
{
    int f = 0;
    for (int d = 0; d < R - 1; ++d) f = f * N + c[d];
    return f;
}

// The x ghost of ghost_value read from the side buffer where the neighbour's
// touching layer holds it (same level: copy; coarser: injection); a finer
// neighbour's mean needs two layers and is read from the interiors.
template <int R, int N>
__host__ __device__ inline double xface_ghost(const double* __restrict__ u, const double* __restrict__ xf,
                                              const face_link& L, uint32_t meta, int dir, const int* c)
{
    using S        = patch_shape<R, N>;
    const int side = (dir & 1) ? 0 : 1; // -x reads the neighbour's x = N-1 layer
    switch (rel_of(meta))
    {
        case REL_SAME:
            return xf[(static_cast<size_t>(L.ids[0]) * 2 + side) * S::FACE + xface_index<R, N>(c)];
        case REL_COARSER:
        {
            int cc[R];
            for (int d = 0; d < R - 1; ++d) cc[d] = static_cast<int>(quad_of(meta, d)) * S::HALF + c[d] / 2;
            return xf[(static_cast<size_t>(L.ids[0]) * 2 + side) * S::FACE + xface_index<R, N>(cc)];
        }
        default: return ghost_value<R, N>(u, L, meta, dir, c);
    }
}

// Builds the x-face side buffers of a state from its interiors.
template <int R, int N, int NT>
__global__ __launch_bounds__(NT) void k_xfaces(const double* __restrict__ u, double* __restrict__ xf, uint32_t nleaves)
{
    using S             = patch_shape<R, N>;
    const uint32_t leaf = blockIdx.x;
    if (leaf >= nleaves) return;
    for (int t = threadIdx.x; t < 2 * S::FACE; t += NT)
    {
        const int side = t / S::FACE;
        int       c[R], f = t % S::FACE;
        for (int d = R - 2; d >= 0; --d)
        {
            c[d] = f % N;
            f /= N;
        }
        c[R - 1] = side ? N - 1 : 0;
        xf[static_cast<size_t>(leaf) * 2 * S::FACE + t] = u[static_cast<size_t>(leaf) * S::NI + S::lin(c)];
    }
}

// Index into u of the one interior cell that gives the ghost value of cell c
// when the direction-`dir` neighbour (slot nb) is same-level (copy) or
// coarser (injection): ghost_value's single-cell rules.
template <int R, int N>
__host__ __device__ inline size_t ghost_source(int32_t nb, uint32_t meta, int dir, const int* c)
{
    using S           = patch_shape<R, N>;
    const int  dim    = dir >> 1;
    const int  nrm    = (dir & 1) ? c[dim] - N : c[dim] + N;
    const bool coarse = rel_of(meta) == REL_COARSER;
    int        cc[R];
    for (int d = 0; d < R; ++d)
    {
        const int v = (d == dim) ? nrm : c[d];
        cc[d]       = coarse ? static_cast<int>(quad_of(meta, d)) * S::HALF + v / 2 : v;
    }
    return static_cast<size_t>(nb) * S::NI + S::lin(cc);
}

// The same cell in the x-face side buffer (dir = -x or +x).
template <int R, int N>
__host__ __device__ inline size_t xface_source(int32_t nb, uint32_t meta, int dir, const int* c)
{
    using S           = patch_shape<R, N>;
    const int  side   = (dir & 1) ? 0 : 1;
    const bool coarse = rel_of(meta) == REL_COARSER;
    int        cc[R];
    for (int d = 0; d < R - 1; ++d) cc[d] = coarse ? static_cast<int>(quad_of(meta, d)) * S::HALF + c[d] / 2 : c[d];
    return (static_cast<size_t>(nb) * 2 + side) * S::FACE + xface_index<R, N>(cc);
}

// A leaf's face records: relation/quadrant word and first neighbour slot per
// direction, read with one batch of scalar loads.
template <int R>
__device__ inline void load_faces(const face_link* __restrict__ L, const uint32_t* __restrict__ M, uint32_t* m,
                                  int32_t* nb)
{
#pragma unroll
    for (int k = 0; k < 2 * R; ++k)
    {
        m[k]  = M[k];
        nb[k] = L[k].ids[0];
    }
}

// The face records of leaf i of a uniform periodic tree whose leaves are all
// at level ul (so the leaves are the 2^(R ul) Morton codes in order and slot
// = Morton code): every neighbour is same-level, its slot is the Morton code
// of the neighbouring coordinates.  No table is read.
template <int R>
__host__ __device__ inline void uniform_faces(uint32_t i, int ul, uint32_t* m, int32_t* nb)
{
    const uint32_t mask = (1u << ul) - 1u;
#pragma unroll
    for (int k = 0; k < 2 * R; ++k)
    {
        const int a = R - 1 - (k >> 1); // physical axis of layout dim k / 2 (Morton bit a of each group)
        uint32_t  c = 0;
        for (int b = 0; b < ul; ++b) c |= ((i >> (R * b + a)) & 1u) << b;
        c          = (c + ((k & 1) ? 1u : mask)) & mask;
        uint32_t j = i;
        for (int b = 0; b < ul; ++b) j = (j & ~(1u << (R * b + a))) | (((c >> b) & 1u) << (R * b + a));
        m[k]  = REL_SAME;
        nb[k] = static_cast<int32_t>(j);
    }
}

// A leaf's face records and level fetched with one vector load per lane
// (lane k < 2R: relation word k; 2R <= k < 4R: first slot of face k - 2R;
// lane 4R: low word of the key), so that the load counts in the in-order
// vector-memory counter: a record fetched behind the current leaf's loads
// does not hold up the current leaf.  unpack_faces reads them back per lane.
template <int R>
__device__ inline uint32_t fetch_faces(const face_link* __restrict__ L, const uint32_t* __restrict__ M,
                                       const uint64_t* __restrict__ key, int lane)
{
    if (lane < 2 * R) return M[lane];
    if (lane < 4 * R) return static_cast<uint32_t>(L[lane - 2 * R].ids[0]);
    return static_cast<uint32_t>(*key);
}

template <int R>
__device__ inline int unpack_faces(uint32_t rec, uint32_t* m, int32_t* nb)
{
#pragma unroll
    for (int k = 0; k < 2 * R; ++k)
    {
        m[k]  = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(rec), k));
        nb[k] = __builtin_amdgcn_readlane(static_cast<int>(rec), 2 * R + k);
    }
    return __builtin_amdgcn_readlane(static_cast<int>(rec), 4 * R) & 0x3F;
}

// Whether none of a leaf's faces has finer neighbours (every ghost is then a
// single cell and the gathers below issue all their loads back to back).
template <int R>
__device__ inline bool single_cell_ghosts(const uint32_t* m)
{
    uint32_t finer = 0;
#pragma unroll
    for (int k = 0; k < 2 * R; ++k) finer |= (rel_of(m[k]) == REL_FINER) ? 1u : 0u;
    return finer == 0;
}

// The six ghost values of a lane of the 3D wave layout (lane = 8 hi + lo),
// from the leaf's face records (m, nb; L, M for the finer-neighbour path):
// g[0], g[1] = -z / +z of the column (y, x) = (hi, lo); g[2], g[3] = the
// (z, x) = (hi, lo) cell of the -y / +y face; g[4], g[5] = the (z, y) =
// (hi, lo) cell of the -x / +x face (from the side buffer with XF).  The
// leaf's table entries are read first and, unless a face has finer
// neighbours, all six loads are in flight at once.
template <bool XF>
__device__ inline void gather_ghosts_3d8(const double* __restrict__ u, const double* __restrict__ xf,
                                         const uint32_t* m, const int32_t* nb, const face_link* __restrict__ L,
                                         const uint32_t* __restrict__ M, int hi, int lo, bool want_xm, bool want_xp,
                                         double* g)
{
    const int c[6][3] = { { -1, hi, lo }, { 8, hi, lo }, { hi, -1, lo }, { hi, 8, lo }, { hi, lo, -1 }, { hi, lo, 8 } };
    if (single_cell_ghosts<3>(m))
    {
        size_t src[6];
#pragma unroll
        for (int k = 0; k < 6; ++k)
            src[k] = (XF && k >= 4) ? xface_source<3, 8>(nb[k], m[k], k, c[k]) : ghost_source<3, 8>(nb[k], m[k], k, c[k]);
#pragma unroll
        for (int k = 0; k < 4; ++k) g[k] = u[src[k]];
        const double* xb = XF ? xf : u;
        g[4]             = want_xm ? xb[src[4]] : 0.0;
        g[5]             = want_xp ? xb[src[5]] : 0.0;
        return;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k)
    {
        if ((k == 4 && !want_xm) || (k == 5 && !want_xp))
            g[k] = 0.0;
        else if (XF && k >= 4)
            g[k] = xface_ghost<3, 8>(u, xf, L[k], M[k], k, c[k]);
        else
            g[k] = ghost_value<3, 8>(u, L[k], M[k], k, c[k]);
    }
}

// The ghost value a lane of the 2D wave layout gathers: cell t of face q
// (lane = 16 q + t), from the side buffer for the x faces with XF.
template <bool XF>
__device__ inline double gather_ghost_2d16(const double* __restrict__ u, const double* __restrict__ xf,
                                           const uint32_t* ms, const int32_t* nbs, const face_link* __restrict__ L,
                                           const uint32_t* __restrict__ M, int q, int t)
{
    int c[2];
    c[0] = (q == 0) ? -1 : (q == 1) ? 16 : t;
    c[1] = (q == 2) ? -1 : (q == 3) ? 16 : t;
    if (single_cell_ghosts<2>(ms))
    {
        // the lane's own face record, selected from the four uniform ones
        uint32_t m  = ms[0];
        int32_t  nb = nbs[0];
#pragma unroll
        for (int k = 1; k < 4; ++k)
            if (q == k)
            {
                m  = ms[k];
                nb = nbs[k];
            }
        if (XF && q >= 2) return xf[xface_source<2, 16>(nb, m, q, c)];
        return u[ghost_source<2, 16>(nb, m, q, c)];
    }
    if (XF && q >= 2) return xface_ghost<2, 16>(u, xf, L[q], M[q], q, c);
    return ghost_value<2, 16>(u, L[q], M[q], q, c);
}

// Upwind update of one cell from its neighbours along each physical axis d
// (lo[d] = u_L, hi[d] = u_R): the expression order of
// amr_solver::time_step_cpu (amr_solver.hpp:302-347) and
// AdvectionPhysics::rusanovFlux (AdvectionPhysics.hpp:45-66):
//   total = 0; total -= dt/dx_d * (F(u, u_R) - F(u_L, u)) for d = x, y(, z);
//   F(a, b) = 0.5 * (a v + b v) - (0.5 |v|) (b - a); out = u + total.
__host__ __device__ inline double upwind_flux(double a, double b, double v, double hs)
{
    return 0.5 * (a * v + b * v) - hs * (b - a);
}

// The update from the cell's face fluxes, fL[d] = F(u_L, u), fR[d] = F(u, u_R).
// A face flux is the same bits whether computed for the cell on its left or
// on its right, so kernels may evaluate each face once.
template <int R>
__host__ __device__ inline double update_from_fluxes(double uc, const double* fL, const double* fR, const double* dtdx)
{
    double total = 0.0;
    for (int d = 0; d < R; ++d) total -= dtdx[d] * (fR[d] - fL[d]);
    return uc + total;
}

template <int R>
__host__ __device__ inline double upwind_update(double uc, const double* lo, const double* hi, const double* dtdx,
                                                const double* vel, const double* hs)
{
    double fL[R], fR[R];
    for (int d = 0; d < R; ++d)
    {
        fL[d] = upwind_flux(lo[d], uc, vel[d], hs[d]);
        fR[d] = upwind_flux(uc, hi[d], vel[d], hs[d]);
    }
    return update_from_fluxes<R>(uc, fL, fR, dtdx);
}

// Upwind update of interior cell i from the staged tile.
template <int R, int N>
__host__ __device__ inline double upwind_cell(const double* tile, int i, const double* dtdx, const double* vel,
                                              const double* hs)
{
    using T = tile_shape<R, N>;
    int c[R];
    patch_shape<R, N>::coords(i, c);
    const int t = T::idx(c);
    double    lo[R], hi[R];
    for (int d = 0; d < R; ++d)
    {
        const int s = T::stride(R - 1 - d); // physical axis d
        lo[d]       = tile[t - s];
        hi[d]       = tile[t + s];
    }
    return upwind_update<R>(tile[t], lo, hi, dtdx, vel, hs);
}

// max over layout dims of max(|u[c+s]-u[c]|, |u[c]-u[c-s]|) for cell i.
template <int R, int N>
__host__ __device__ inline double gradient_cell(const double* tile, int i)
{
    using T = tile_shape<R, N>;
    int c[R];
    patch_shape<R, N>::coords(i, c);
    const int t = T::idx(c);
    double    g = 0.0;
    for (int j = 0; j < R; ++j)
    {
        const int s = T::stride(j);
        g           = fmax(g, fmax(fabs(tile[t + s] - tile[t]), fabs(tile[t] - tile[t - s])));
    }
    return g;
}

// Cell i of the padded view ((N+2H)^R, row-major): interior copy, face halo
// gathered; *kind = 0 interior, 1 face halo, 2 edge/corner (value 0).
template <int R, int N, int H>
__host__ __device__ inline double padded_value(const double* __restrict__ u, uint32_t leaf,
                                               const face_link* __restrict__ links,
                                               const uint32_t* __restrict__ metas, int i, int* kind)
{
    using S         = patch_shape<R, N>;
    constexpr int P = N + 2 * H;
    int c[R], r = i, outside = 0, odim = 0;
    for (int d = R - 1; d >= 0; --d)
    {
        c[d] = r % P - H;
        r /= P;
        if (c[d] < 0 || c[d] >= N)
        {
            ++outside;
            odim = d;
        }
    }
    *kind = outside < 2 ? outside : 2;
    if (outside == 0) return u[static_cast<size_t>(leaf) * S::NI + S::lin(c)];
    if (outside == 1)
    {
        const int    dir = 2 * odim + (c[odim] >= N ? 1 : 0);
        const size_t e   = static_cast<size_t>(leaf) * S::NDIR + dir;
        return ghost_value<R, N>(u, links[e], metas[e], dir, c);
    }
    return 0.0;
}

__host__ __device__ inline int8_t decide(double value, int level, double rthr, double cthr, int min_level,
                                         int max_level)
{
    if (value > rthr && level < max_level) return 1;
    if (value < cthr && level > min_level) return 2;
    return 0;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

template <int R, int N, int NT>
__device__ inline void stage_tile(double* tile, const double* __restrict__ u, uint32_t leaf,
                                  const face_link* __restrict__ links, const uint32_t* __restrict__ metas)
{
    using S          = patch_shape<R, N>;
    const double* up = u + static_cast<size_t>(leaf) * S::NI;
    for (int i = threadIdx.x; i < S::NI; i += NT) stage_interior<R, N>(tile, up, i);
    for (int g = threadIdx.x; g < S::NDIR * S::FACE; g += NT) stage_ghost<R, N>(tile, u, leaf, links, metas, g);
}

// One time step for every leaf; one workgroup per leaf patch
// (XCD-contiguous ranges).
template <int R, int N, int NT>
__global__ __launch_bounds__(NT) void k_step(const double* __restrict__ u, double* __restrict__ un,
                                             const uint64_t* __restrict__ keys,
                                             const face_link* __restrict__ links,
                                             const uint32_t* __restrict__ metas, step_params prm,
                                             const step_params* __restrict__ table,
                                             const uint16_t* __restrict__ inst, uint32_t nleaves)
{
    using S = patch_shape<R, N>;
    __shared__ double tile[tile_shape<R, N>::TS];
    const uint32_t leaf = sweep_item(blockIdx.x, gridDim.x, nleaves, prm.reverse);
    if (leaf >= nleaves) return;
    stage_tile<R, N, NT>(tile, u, leaf, links, metas);
    __syncthreads();
    const step_params& P  = params_for(prm, table, inst, leaf);
    const int          lv = key_level(keys[leaf]);
    double             dtdx[R], vel[R], hs[R];
    for (int d = 0; d < R; ++d)
    {
        dtdx[d] = P.dtdx[lv][d];
        vel[d]  = P.vel[d];
        hs[d]   = P.hs[d];
    }
    double* out = un + static_cast<size_t>(leaf) * S::NI;
    for (int i = threadIdx.x; i < S::NI; i += NT) out[i] = upwind_cell<R, N>(tile, i, dtdx, vel, hs);
}

// 3D 8^3 patches, one wave (64 lanes) per leaf.  Lane l owns the z-column at
// (y, x) = (l / 8, l % 8) in registers: x- and y-neighbours come from lanes
// l -+ 1 and l -+ 8, z-neighbours from the lane's own registers.  The six
// ghost faces are gathered one face cell per lane; the -+z faces land in the
// lane that needs them, the -+y / -+x faces go through 2 KiB of LDS.  With
// no interior tile in LDS, 32 leaves are in flight per CU.  With XF the x
// ghosts come from the x-face side buffers (xf_in) and the kernel writes
// those of the new state (xf_out).
template <bool NT, bool XF = false>
__global__ __launch_bounds__(64) void k_step_3d8(const double* __restrict__ u, double* __restrict__ un,
                                                 const uint64_t* __restrict__ keys,
                                                 const face_link* __restrict__ links,
                                                 const uint32_t* __restrict__ metas, step_params prm,
                                                 const step_params* __restrict__ table,
                                                 const uint16_t* __restrict__ inst, uint32_t nleaves,
                                                 const double* __restrict__ xf_in = nullptr,
                                                 double* __restrict__ xf_out = nullptr, int ulevel = -1)
{
    constexpr int NI = 512;
    __shared__ double gface[4][64]; // -y, +y (index z*8+x), -x, +x (index z*8+y)
    __shared__ double xs[XF ? 2 : 1][64];
    const uint32_t leaf = sweep_item(blockIdx.x, gridDim.x, nleaves, prm.reverse);
    if (leaf >= nleaves) return;
    const int     lane = threadIdx.x, hi3 = lane >> 3, lo3 = lane & 7;
    const double* up   = u + static_cast<size_t>(leaf) * NI;
    double        r[8];
#pragma unroll
    for (int z = 0; z < 8; ++z) r[z] = up[z * 64 + lane];
    const face_link* L = links + static_cast<size_t>(leaf) * 6;
    const uint32_t*  M = metas + static_cast<size_t>(leaf) * 6;
    uint32_t         m[6];
    int32_t          nb[6];
    if (ulevel >= 0)
        uniform_faces<3>(leaf, ulevel, m, nb);
    else
        load_faces<3>(L, M, m, nb);
    double g[6];
    gather_ghosts_3d8<XF>(u, xf_in, m, nb, L, M, hi3, lo3, true, true, g);
    const double gzm = g[0], gzp = g[1];
#pragma unroll
    for (int k = 0; k < 4; ++k) gface[k][lane] = g[2 + k];
    __syncthreads();
    const step_params& P       = params_for(prm, table, inst, leaf);
    const int          lv      = ulevel >= 0 ? ulevel : key_level(keys[leaf]);
    const double       dtdx[3] = { P.dtdx[lv][0], P.dtdx[lv][1], P.dtdx[lv][2] };
    const int    y = hi3, x = lo3;
    double*      out = un + static_cast<size_t>(leaf) * NI;
    // z faces are evaluated once: the +z flux of cell z is the -z flux of z+1
    double fz = upwind_flux(gzm, r[0], P.vel[2], P.hs[2]);
#pragma unroll
    for (int z = 0; z < 8; ++z)
    {
        const double uc = r[z];
        const double xl = __shfl_up(uc, 1, 64);
        const double xr = __shfl_down(uc, 1, 64);
        const double yl = __shfl_up(uc, 8, 64);
        const double yr = __shfl_down(uc, 8, 64);
        const double lx = (x == 0) ? gface[2][z * 8 + y] : xl;
        const double hx = (x == 7) ? gface[3][z * 8 + y] : xr;
        const double ly = (y == 0) ? gface[0][z * 8 + x] : yl;
        const double hy = (y == 7) ? gface[1][z * 8 + x] : yr;
        const double hz = (z == 7) ? gzp : r[z < 7 ? z + 1 : 7];
        const double fzr = upwind_flux(uc, hz, P.vel[2], P.hs[2]);
        const double fL[3] = { upwind_flux(lx, uc, P.vel[0], P.hs[0]), upwind_flux(ly, uc, P.vel[1], P.hs[1]), fz };
        const double fR[3] = { upwind_flux(uc, hx, P.vel[0], P.hs[0]), upwind_flux(uc, hy, P.vel[1], P.hs[1]), fzr };
        const double o     = update_from_fluxes<3>(uc, fL, fR, dtdx);
        store_cell<NT>(out + z * 64 + lane, o);
        if constexpr (XF)
        {
            if (x == 0) xs[0][z * 8 + y] = o;
            if (x == 7) xs[1][z * 8 + y] = o;
        }
        fz = fzr;
    }
    if constexpr (XF)
    {
        __syncthreads();
        xf_out[static_cast<size_t>(leaf) * 128 + lane]      = xs[0][lane];
        xf_out[static_cast<size_t>(leaf) * 128 + 64 + lane] = xs[1][lane];
    }
}

// 3D 8^3 patches, one wave per K consecutive leaves, k_step_3d8's lane
// layout.  The table records of the next leaf are fetched behind the current
// leaf's interior and ghost loads, so after the first leaf a leaf's ghost
// loads are issued together with its interior loads instead of one table
// round trip later.
template <int K>
__global__ __launch_bounds__(64) void k_step_3d8_pf(const double* __restrict__ u, double* __restrict__ un,
                                                    const uint64_t* __restrict__ keys,
                                                    const face_link* __restrict__ links,
                                                    const uint32_t* __restrict__ metas, step_params prm,
                                                    const step_params* __restrict__ table,
                                                    const uint16_t* __restrict__ inst, uint32_t nleaves)
{
    constexpr int NI = 512;
    __shared__ double gface[4][64]; // -y, +y (index z*8+x), -x, +x (index z*8+y)
    const uint32_t nw = (nleaves + K - 1) / K;
    const uint32_t w  = sweep_item(blockIdx.x, gridDim.x, nw, prm.reverse);
    if (w >= nw) return;
    const int      lane = threadIdx.x, hi3 = lane >> 3, lo3 = lane & 7;
    const int      y = hi3, x = lo3;
    const uint32_t first = w * K, end = (first + K < nleaves) ? first + K : nleaves;
    uint32_t       rec   = fetch_faces<3>(links + static_cast<size_t>(first) * 6, metas + static_cast<size_t>(first) * 6,
                                          keys + first, lane);
#pragma unroll 1
    for (uint32_t leaf = first; leaf < end; ++leaf)
    {
        uint32_t  m[6];
        int32_t   nb[6];
        const int lv = unpack_faces<3>(rec, m, nb);
        const double* up = u + static_cast<size_t>(leaf) * NI;
        double        r[8];
#pragma unroll
        for (int z = 0; z < 8; ++z) r[z] = up[z * 64 + lane];
        double g[6];
        gather_ghosts_3d8<false>(u, nullptr, m, nb, links + static_cast<size_t>(leaf) * 6,
                                 metas + static_cast<size_t>(leaf) * 6, hi3, lo3, true, true, g);
        if (leaf + 1 < end)
            rec = fetch_faces<3>(links + static_cast<size_t>(leaf + 1) * 6, metas + static_cast<size_t>(leaf + 1) * 6,
                                 keys + leaf + 1, lane);
        const double gzm = g[0], gzp = g[1];
#pragma unroll
        for (int k = 0; k < 4; ++k) gface[k][lane] = g[2 + k];
        __syncthreads();
        const step_params& P       = params_for(prm, table, inst, leaf);
        const double       dtdx[3] = { P.dtdx[lv][0], P.dtdx[lv][1], P.dtdx[lv][2] };
        double*            out     = un + static_cast<size_t>(leaf) * NI;
        double             fz      = upwind_flux(gzm, r[0], P.vel[2], P.hs[2]);
#pragma unroll
        for (int z = 0; z < 8; ++z)
        {
            const double uc = r[z];
            const double xl = __shfl_up(uc, 1, 64);
            const double xr = __shfl_down(uc, 1, 64);
            const double yl = __shfl_up(uc, 8, 64);
            const double yr = __shfl_down(uc, 8, 64);
            const double lx = (x == 0) ? gface[2][z * 8 + y] : xl;
            const double hx = (x == 7) ? gface[3][z * 8 + y] : xr;
            const double ly = (y == 0) ? gface[0][z * 8 + x] : yl;
            const double hy = (y == 7) ? gface[1][z * 8 + x] : yr;
            const double hz = (z == 7) ? gzp : r[z < 7 ? z + 1 : 7];
            const double fzr = upwind_flux(uc, hz, P.vel[2], P.hs[2]);
            const double fL[3] = { upwind_flux(lx, uc, P.vel[0], P.hs[0]), upwind_flux(ly, uc, P.vel[1], P.hs[1]), fz };
            const double fR[3] = { upwind_flux(uc, hx, P.vel[0], P.hs[0]), upwind_flux(uc, hy, P.vel[1], P.hs[1]), fzr };
            out[z * 64 + lane] = update_from_fluxes<3>(uc, fL, fR, dtdx);
            fz                 = fzr;
        }
        __syncthreads(); // gface is rewritten by the next leaf
    }
}

// 3D 8^3 patches, one wave per run of same-level leaves along +x (leaf_run).
// The lane layout is k_step_3d8's.  Walking the run front to back, the -x
// ghosts of a leaf are the previous leaf's x = 7 column (left in LDS) and
// its +x ghosts the next leaf's x = 0 column (that leaf's interior is
// loaded before the current leaf is updated, so its loads overlap the
// update); only the run's two end faces are gathered from memory.  An x face
// touches every 64 B row of the neighbour patch, so this removes most of the
// ghost traffic that does not hit in L2.
template <bool NT>
__global__ __launch_bounds__(64) void k_step_3d8_run(const double* __restrict__ u, double* __restrict__ un,
                                                     const uint64_t* __restrict__ keys,
                                                     const face_link* __restrict__ links,
                                                     const uint32_t* __restrict__ metas, step_params prm,
                                                     const step_params* __restrict__ table,
                                                     const uint16_t* __restrict__ inst,
                                                     const leaf_run* __restrict__ runs, uint32_t nruns)
{
    constexpr int NI = 512;
    __shared__ double gface[4][64]; // -y, +y (index z*8+x), -x, +x (index z*8+y)
    const uint32_t r = sweep_item(blockIdx.x, gridDim.x, nruns, prm.reverse);
    if (r >= nruns) return;
    const int      lane = threadIdx.x, hi3 = lane >> 3, lo3 = lane & 7;
    const int      y = hi3, x = lo3;
    const leaf_run run  = runs[r];
    uint32_t       leaf = run.head;
    double         cur[8], nxt[8];
    {
        const double* up = u + static_cast<size_t>(leaf) * NI;
#pragma unroll
        for (int z = 0; z < 8; ++z)
        {
            cur[z] = up[z * 64 + lane];
            nxt[z] = 0.0;
        }
    }
    for (uint32_t i = 0; i < run.len; ++i)
    {
        const face_link* L    = links + static_cast<size_t>(leaf) * 6;
        const uint32_t*  M    = metas + static_cast<size_t>(leaf) * 6;
        const bool       more = i + 1 < run.len;
        // ghosts of this leaf first, then the next leaf's interior, so that
        // the update below waits only for the ghosts
        uint32_t m[6];
        int32_t  nb[6];
        load_faces<3>(L, M, m, nb);
        double g[6];
        gather_ghosts_3d8<false>(u, nullptr, m, nb, L, M, hi3, lo3, i == 0, !more, g);
        const double gzm = g[0], gzp = g[1];
        gface[0][lane]   = g[2];
        gface[1][lane]   = g[3];
        if (i == 0) gface[2][lane] = g[4];
        if (!more) gface[3][lane] = g[5];
        const uint32_t next = more ? static_cast<uint32_t>(L[5].ids[0]) : leaf;
        if (more)
        {
            const double* np = u + static_cast<size_t>(next) * NI;
#pragma unroll
            for (int z = 0; z < 8; ++z) nxt[z] = np[z * 64 + lane];
        }
        __syncthreads();
        const step_params& P       = params_for(prm, table, inst, leaf);
        const int          lv      = key_level(keys[leaf]);
        const double       dtdx[3] = { P.dtdx[lv][0], P.dtdx[lv][1], P.dtdx[lv][2] };
        double*            out     = un + static_cast<size_t>(leaf) * NI;
        double             fz      = upwind_flux(gzm, cur[0], P.vel[2], P.hs[2]); // z faces once
#pragma unroll
        for (int z = 0; z < 8; ++z)
        {
            const double uc = cur[z];
            const double xl = __shfl_up(uc, 1, 64);
            const double xr = __shfl_down(uc, 1, 64);
            const double yl = __shfl_up(uc, 8, 64);
            const double yr = __shfl_down(uc, 8, 64);
            const double xn = __shfl(nxt[z], lane & ~7, 64); // next leaf, x = 0, same (z, y)
            const double lx = (x == 0) ? gface[2][z * 8 + y] : xl;
            const double hx = (x == 7) ? (more ? xn : gface[3][z * 8 + y]) : xr;
            const double ly = (y == 0) ? gface[0][z * 8 + x] : yl;
            const double hy = (y == 7) ? gface[1][z * 8 + x] : yr;
            const double hz = (z == 7) ? gzp : cur[z < 7 ? z + 1 : 7];
            const double fzr = upwind_flux(uc, hz, P.vel[2], P.hs[2]);
            const double fL[3] = { upwind_flux(lx, uc, P.vel[0], P.hs[0]), upwind_flux(ly, uc, P.vel[1], P.hs[1]), fz };
            const double fR[3] = { upwind_flux(uc, hx, P.vel[0], P.hs[0]), upwind_flux(uc, hy, P.vel[1], P.hs[1]), fzr };
            store_cell<NT>(out + z * 64 + lane, update_from_fluxes<3>(uc, fL, fR, dtdx));
            fz = fzr;
        }
        __syncthreads();
        // this leaf's x = 7 column is the next leaf's -x ghost face
#pragma unroll
        for (int z = 0; z < 8; ++z)
        {
            if (x == 7) gface[2][z * 8 + y] = cur[z];
            cur[z] = nxt[z];
        }
        leaf = next;
        __syncthreads();
    }
}

// 3D 8^3 patches, one workgroup of 8 waves per group of 8 consecutive
// leaves (a sibling octet wherever the tree is locally uniform), one wave per
// leaf as in k_step_3d8.  Every wave publishes its six boundary faces in LDS;
// a ghost face whose neighbour is a same-level leaf of the same group is read
// from there, and only the faces leaving the group are gathered from HBM/L2.
// In a uniform octree that halves the global ghost traffic.
constexpr int kGroup3d = 8;

template <bool NT>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_step_3d8_group(
                                                        const double* __restrict__ u, double* __restrict__ un,
                                                        const uint64_t* __restrict__ keys,
                                                        const face_link* __restrict__ links,
                                                        const uint32_t* __restrict__ metas, step_params prm,
                                                        const step_params* __restrict__ table,
                                                        const uint16_t* __restrict__ inst, uint32_t nleaves)
{
    constexpr int NI = 512;
    // own boundary faces, [wave][dir][cell]: dirs 0/1 = z=0/z=7 planes (cell
    // y*8+x), 2/3 = y=0/y=7 (cell z*8+x), 4/5 = x=0/x=7 (cell z*8+y)
    __shared__ double faces[kGroup3d][6][64];
    // gathered ghost faces for the -+y / -+x directions, same cell indexing
    __shared__ double gbuf[kGroup3d][4][64];
    const uint32_t ngroups = (nleaves + kGroup3d - 1) / kGroup3d;
    const uint32_t group   = xcd_item(blockIdx.x, gridDim.x);
    if (group >= ngroups) return; // uniform over the workgroup
    const int      w      = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t first  = group * kGroup3d;
    const uint32_t leaf   = first + w;
    const bool     active = leaf < nleaves;
    const int      y = lane >> 3, x = lane & 7;
    double         r[8];
    double         gzm = 0.0, gzp = 0.0;
    int            nb_wave[6]; // -1: gathered, else the group wave holding the neighbour
    if (active)
    {
        const double* up = u + static_cast<size_t>(leaf) * NI;
#pragma unroll
        for (int z = 0; z < 8; ++z) r[z] = up[z * 64 + lane];
        faces[w][0][lane] = r[0];
        faces[w][1][lane] = r[7];
#pragma unroll
        for (int z = 0; z < 8; ++z)
        {
            if (y == 0) faces[w][2][z * 8 + x] = r[z];
            if (y == 7) faces[w][3][z * 8 + x] = r[z];
            if (x == 0) faces[w][4][z * 8 + y] = r[z];
            if (x == 7) faces[w][5][z * 8 + y] = r[z];
        }
        const face_link* L = links + static_cast<size_t>(leaf) * 6;
        const uint32_t*  M = metas + static_cast<size_t>(leaf) * 6;
#pragma unroll
        for (int dir = 0; dir < 6; ++dir)
        {
            const int32_t nb = L[dir].ids[0];
            nb_wave[dir]     = (rel_of(M[dir]) == REL_SAME && nb >= static_cast<int32_t>(first) &&
                            nb < static_cast<int32_t>(first + kGroup3d))
                                   ? nb - static_cast<int32_t>(first)
                                   : -1;
        }
        int c[3];
        if (nb_wave[0] < 0)
        {
            c[0] = -1, c[1] = y, c[2] = x;
            gzm  = ghost_value<3, 8>(u, L[0], M[0], 0, c);
        }
        if (nb_wave[1] < 0)
        {
            c[0] = 8, c[1] = y, c[2] = x;
            gzp  = ghost_value<3, 8>(u, L[1], M[1], 1, c);
        }
        // -+y faces: face cell (z, x) = (lane / 8, lane % 8)
        if (nb_wave[2] < 0)
        {
            c[0] = y, c[1] = -1, c[2] = x;
            gbuf[w][0][lane] = ghost_value<3, 8>(u, L[2], M[2], 2, c);
        }
        if (nb_wave[3] < 0)
        {
            c[0] = y, c[1] = 8, c[2] = x;
            gbuf[w][1][lane] = ghost_value<3, 8>(u, L[3], M[3], 3, c);
        }
        // -+x faces: face cell (z, y) = (lane / 8, lane % 8)
        if (nb_wave[4] < 0)
        {
            c[0] = y, c[1] = x, c[2] = -1;
            gbuf[w][2][lane] = ghost_value<3, 8>(u, L[4], M[4], 4, c);
        }
        if (nb_wave[5] < 0)
        {
            c[0] = y, c[1] = x, c[2] = 8;
            gbuf[w][3][lane] = ghost_value<3, 8>(u, L[5], M[5], 5, c);
        }
    }
    __syncthreads();
    if (!active) return;
    // a ghost of direction dir is the neighbour's face on the opposite side
    if (nb_wave[0] >= 0) gzm = faces[nb_wave[0]][1][lane];
    if (nb_wave[1] >= 0) gzp = faces[nb_wave[1]][0][lane];
    const double* gym = nb_wave[2] >= 0 ? faces[nb_wave[2]][3] : gbuf[w][0];
    const double* gyp = nb_wave[3] >= 0 ? faces[nb_wave[3]][2] : gbuf[w][1];
    const double* gxm = nb_wave[4] >= 0 ? faces[nb_wave[4]][5] : gbuf[w][2];
    const double* gxp = nb_wave[5] >= 0 ? faces[nb_wave[5]][4] : gbuf[w][3];
    const step_params& P       = params_for(prm, table, inst, leaf);
    const int          lv      = key_level(keys[leaf]);
    const double       dtdx[3] = { P.dtdx[lv][0], P.dtdx[lv][1], P.dtdx[lv][2] };
    double*       out     = un + static_cast<size_t>(leaf) * NI;
#pragma unroll
    for (int z = 0; z < 8; ++z)
    {
        const double uc = r[z];
        const double xl = __shfl_up(uc, 1, 64);
        const double xr = __shfl_down(uc, 1, 64);
        const double yl = __shfl_up(uc, 8, 64);
        const double yr = __shfl_down(uc, 8, 64);
        double       lo[3], hi[3];
        lo[0] = (x == 0) ? gxm[z * 8 + y] : xl;
        hi[0] = (x == 7) ? gxp[z * 8 + y] : xr;
        lo[1] = (y == 0) ? gym[z * 8 + x] : yl;
        hi[1] = (y == 7) ? gyp[z * 8 + x] : yr;
        if (z == 0)
            lo[2] = gzm;
        else
            lo[2] = r[z > 0 ? z - 1 : 0];
        if (z == 7)
            hi[2] = gzp;
        else
            hi[2] = r[z < 7 ? z + 1 : 7];
        store_cell<NT>(out + z * 64 + lane, upwind_update<3>(uc, lo, hi, dtdx, P.vel, P.hs));
    }
}

// 2D 16^2 patches, one wave per leaf.  Lane l owns x = l % 16 of rows
// y = 4 (l / 16) .. 4 (l / 16) + 3; x-neighbours from lanes l -+ 1, the row
// above / below a lane's four from lanes l -+ 16; the four ghost faces are
// gathered one face cell per lane into 512 B of LDS.
template <bool NT, bool XF = false>
__global__ __launch_bounds__(64) void k_step_2d16(const double* __restrict__ u, double* __restrict__ un,
                                                  const uint64_t* __restrict__ keys,
                                                  const face_link* __restrict__ links,
                                                  const uint32_t* __restrict__ metas, step_params prm,
                                                  const step_params* __restrict__ table,
                                                  const uint16_t* __restrict__ inst, uint32_t nleaves,
                                                  const double* __restrict__ xf_in = nullptr,
                                                  double* __restrict__ xf_out = nullptr, int ulevel = -1)
{
    constexpr int NI = 256;
    __shared__ double gface[4][16]; // -y, +y (index x), -x, +x (index y)
    __shared__ double xs[XF ? 32 : 1];  // new x = 0 / x = 15 columns
    const uint32_t leaf = sweep_item(blockIdx.x, gridDim.x, nleaves, prm.reverse);
    if (leaf >= nleaves) return;
    const int     lane = threadIdx.x, x = lane & 15, q = lane >> 4;
    const double* up   = u + static_cast<size_t>(leaf) * NI;
    double        r[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) r[j] = up[(4 * q + j) * 16 + x];
    const face_link* L = links + static_cast<size_t>(leaf) * 4;
    const uint32_t*  M = metas + static_cast<size_t>(leaf) * 4;
    uint32_t         m[4];
    int32_t          nb[4];
    if (ulevel >= 0)
        uniform_faces<2>(leaf, ulevel, m, nb);
    else
        load_faces<2>(L, M, m, nb);
    // lane (q, x) gathers cell x of face q
    gface[q][x] = gather_ghost_2d16<XF>(u, xf_in, m, nb, L, M, q, x);
    __syncthreads();
    const step_params& P       = params_for(prm, table, inst, leaf);
    const int          lv      = ulevel >= 0 ? ulevel : key_level(keys[leaf]);
    const double       dtdx[2] = { P.dtdx[lv][0], P.dtdx[lv][1] };
    const double above   = __shfl_up(r[3], 16, 64);  // row 4q - 1
    const double below   = __shfl_down(r[0], 16, 64); // row 4q + 4
    double*      out     = un + static_cast<size_t>(leaf) * NI;
    // y faces are evaluated once: the +y flux of row j is the -y flux of j+1
    double fy = upwind_flux((q == 0) ? gface[0][x] : above, r[0], P.vel[1], P.hs[1]);
#pragma unroll
    for (int j = 0; j < 4; ++j)
    {
        const int    y  = 4 * q + j;
        const double uc = r[j];
        const double xl = __shfl_up(uc, 1, 64);
        const double xr = __shfl_down(uc, 1, 64);
        const double lx = (x == 0) ? gface[2][y] : xl;
        const double hx = (x == 15) ? gface[3][y] : xr;
        const double hy = (j < 3) ? r[j < 3 ? j + 1 : 3] : ((q == 3) ? gface[1][x] : below);
        const double fyr   = upwind_flux(uc, hy, P.vel[1], P.hs[1]);
        const double fL[2] = { upwind_flux(lx, uc, P.vel[0], P.hs[0]), fy };
        const double fR[2] = { upwind_flux(uc, hx, P.vel[0], P.hs[0]), fyr };
        const double o     = update_from_fluxes<2>(uc, fL, fR, dtdx);
        store_cell<NT>(out + y * 16 + x, o);
        if constexpr (XF)
        {
            if (x == 0) xs[y] = o;
            if (x == 15) xs[16 + y] = o;
        }
        fy = fyr;
    }
    if constexpr (XF)
    {
        __syncthreads();
        if (lane < 32) xf_out[static_cast<size_t>(leaf) * 32 + lane] = xs[lane];
    }
}

// 2D 16^2 patches, one wave per K consecutive leaves, k_step_2d16's lane
// layout, the next leaf's table records fetched behind the current leaf's
// loads (as k_step_3d8_pf).
template <int K>
__global__ __launch_bounds__(64) void k_step_2d16_pf(const double* __restrict__ u, double* __restrict__ un,
                                                     const uint64_t* __restrict__ keys,
                                                     const face_link* __restrict__ links,
                                                     const uint32_t* __restrict__ metas, step_params prm,
                                                     const step_params* __restrict__ table,
                                                     const uint16_t* __restrict__ inst, uint32_t nleaves)
{
    constexpr int NI = 256;
    __shared__ double gface[4][16]; // -y, +y (index x), -x, +x (index y)
    const uint32_t nw = (nleaves + K - 1) / K;
    const uint32_t w  = sweep_item(blockIdx.x, gridDim.x, nw, prm.reverse);
    if (w >= nw) return;
    const int      lane = threadIdx.x, x = lane & 15, q = lane >> 4;
    const uint32_t first = w * K, end = (first + K < nleaves) ? first + K : nleaves;
    uint32_t       rec   = fetch_faces<2>(links + static_cast<size_t>(first) * 4, metas + static_cast<size_t>(first) * 4,
                                          keys + first, lane);
#pragma unroll 1
    for (uint32_t leaf = first; leaf < end; ++leaf)
    {
        uint32_t  m[4];
        int32_t   nb[4];
        const int lv = unpack_faces<2>(rec, m, nb);
        const double* up = u + static_cast<size_t>(leaf) * NI;
        double        r[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) r[j] = up[(4 * q + j) * 16 + x];
        const double gq = gather_ghost_2d16<false>(u, nullptr, m, nb, links + static_cast<size_t>(leaf) * 4,
                                                   metas + static_cast<size_t>(leaf) * 4, q, x);
        if (leaf + 1 < end)
            rec = fetch_faces<2>(links + static_cast<size_t>(leaf + 1) * 4, metas + static_cast<size_t>(leaf + 1) * 4,
                                 keys + leaf + 1, lane);
        gface[q][x] = gq;
        __syncthreads();
        const step_params& P       = params_for(prm, table, inst, leaf);
        const double       dtdx[2] = { P.dtdx[lv][0], P.dtdx[lv][1] };
        const double       above   = __shfl_up(r[3], 16, 64);  // row 4q - 1
        const double       below   = __shfl_down(r[0], 16, 64); // row 4q + 4
        double*            out     = un + static_cast<size_t>(leaf) * NI;
        double             fy      = upwind_flux((q == 0) ? gface[0][x] : above, r[0], P.vel[1], P.hs[1]);
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            const int    y  = 4 * q + j;
            const double uc = r[j];
            const double xl = __shfl_up(uc, 1, 64);
            const double xr = __shfl_down(uc, 1, 64);
            const double lx = (x == 0) ? gface[2][y] : xl;
            const double hx = (x == 15) ? gface[3][y] : xr;
            const double hy = (j < 3) ? r[j < 3 ? j + 1 : 3] : ((q == 3) ? gface[1][x] : below);
            const double fyr   = upwind_flux(uc, hy, P.vel[1], P.hs[1]);
            const double fL[2] = { upwind_flux(lx, uc, P.vel[0], P.hs[0]), fy };
            const double fR[2] = { upwind_flux(uc, hx, P.vel[0], P.hs[0]), fyr };
            out[y * 16 + x]    = update_from_fluxes<2>(uc, fL, fR, dtdx);
            fy                 = fyr;
        }
        __syncthreads();
    }
}

// 2D 16^2 patches, one wave per run of same-level leaves along +x, the lane
// layout of k_step_2d16 and the hand-over of k_step_3d8_run: the -x ghosts
// of a leaf are the previous leaf's x = 15 column (left in LDS), its +x
// ghosts the next leaf's x = 0 column (registers of the prefetched leaf).
template <bool NT>
__global__ __launch_bounds__(64) void k_step_2d16_run(const double* __restrict__ u, double* __restrict__ un,
                                                      const uint64_t* __restrict__ keys,
                                                      const face_link* __restrict__ links,
                                                      const uint32_t* __restrict__ metas, step_params prm,
                                                      const step_params* __restrict__ table,
                                                      const uint16_t* __restrict__ inst,
                                                      const leaf_run* __restrict__ runs, uint32_t nruns)
{
    constexpr int NI = 256;
    __shared__ double gface[4][16]; // -y, +y (index x), -x, +x (index y)
    const uint32_t r = sweep_item(blockIdx.x, gridDim.x, nruns, prm.reverse);
    if (r >= nruns) return;
    const int      lane = threadIdx.x, x = lane & 15, q = lane >> 4;
    const leaf_run run  = runs[r];
    uint32_t       leaf = run.head;
    double         cur[4], nxt[4];
    {
        const double* up = u + static_cast<size_t>(leaf) * NI;
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            cur[j] = up[(4 * q + j) * 16 + x];
            nxt[j] = 0.0;
        }
    }
    for (uint32_t i = 0; i < run.len; ++i)
    {
        const bool   more = i + 1 < run.len;
        const size_t e0   = static_cast<size_t>(leaf) * 4;
        // lanes of quarter q gather face q (y faces always, the run's end x faces)
        uint32_t m[4];
        int32_t  nb[4];
        load_faces<2>(links + e0, metas + e0, m, nb);
        if (q < 2 || (q == 2 && i == 0) || (q == 3 && !more))
            gface[q][x] = gather_ghost_2d16<false>(u, nullptr, m, nb, links + e0, metas + e0, q, x);
        const uint32_t next = more ? static_cast<uint32_t>(links[e0 + 3].ids[0]) : leaf;
        if (more)
        {
            const double* np = u + static_cast<size_t>(next) * NI;
#pragma unroll
            for (int j = 0; j < 4; ++j) nxt[j] = np[(4 * q + j) * 16 + x];
        }
        __syncthreads();
        const step_params& P       = params_for(prm, table, inst, leaf);
        const int          lv      = key_level(keys[leaf]);
        const double       dtdx[2] = { P.dtdx[lv][0], P.dtdx[lv][1] };
        const double       above   = __shfl_up(cur[3], 16, 64);  // row 4q - 1
        const double       below   = __shfl_down(cur[0], 16, 64); // row 4q + 4
        double*            out     = un + static_cast<size_t>(leaf) * NI;
        double             fy      = upwind_flux((q == 0) ? gface[0][x] : above, cur[0], P.vel[1], P.hs[1]);
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            const int    y  = 4 * q + j;
            const double uc = cur[j];
            const double xl = __shfl_up(uc, 1, 64);
            const double xr = __shfl_down(uc, 1, 64);
            const double xn = __shfl(nxt[j], lane & ~15, 64); // next leaf, x = 0, same row
            const double lx = (x == 0) ? gface[2][y] : xl;
            const double hx = (x == 15) ? (more ? xn : gface[3][y]) : xr;
            const double hy = (j < 3) ? cur[j < 3 ? j + 1 : 3] : ((q == 3) ? gface[1][x] : below);
            const double fyr   = upwind_flux(uc, hy, P.vel[1], P.hs[1]);
            const double fL[2] = { upwind_flux(lx, uc, P.vel[0], P.hs[0]), fy };
            const double fR[2] = { upwind_flux(uc, hx, P.vel[0], P.hs[0]), fyr };
            store_cell<NT>(out + y * 16 + x, update_from_fluxes<2>(uc, fL, fR, dtdx));
            fy = fyr;
        }
        __syncthreads();
        // this leaf's x = 15 column is the next leaf's -x ghost face
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            if (x == 15) gface[2][4 * q + j] = cur[j];
            cur[j] = nxt[j];
        }
        leaf = next;
        __syncthreads();
    }
}

// Padded view with H ghost layers (the reference's patch after
// halo_exchange_update): interior copied, face halos gathered, corners 0.
template <int R, int N, int H, int NT>
__global__ __launch_bounds__(NT) void k_padded_view(const double* __restrict__ u, double* __restrict__ pad,
                                                    const face_link* __restrict__ links,
                                                    const uint32_t* __restrict__ metas, uint32_t nleaves)
{
    constexpr int  P    = N + 2 * H;
    constexpr int  FLAT = (R == 2) ? P * P : P * P * P;
    const uint32_t leaf = blockIdx.x;
    if (leaf >= nleaves) return;
    for (int i = threadIdx.x; i < FLAT; i += NT)
    {
        int kind;
        pad[static_cast<size_t>(leaf) * FLAT + i] = padded_value<R, N, H>(u, leaf, links, metas, i, &kind);
    }
}

template <int NT>
__device__ inline double block_max(double v, double* red)
{
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1)
    {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    return red[0];
}

// Undivided-gradient criterion: one workgroup per leaf.
template <int R, int N, int NT>
__global__ __launch_bounds__(NT) void k_criterion_gradient(const double* __restrict__ u,
                                                           const uint64_t* __restrict__ keys,
                                                           const face_link* __restrict__ links,
                                                           const uint32_t* __restrict__ metas, double rthr,
                                                           double cthr, int min_level, int max_level,
                                                           int8_t* __restrict__ flags, uint32_t nleaves)
{
    using S = patch_shape<R, N>;
    __shared__ double tile[tile_shape<R, N>::TS];
    __shared__ double red[NT];
    const uint32_t leaf = blockIdx.x;
    if (leaf >= nleaves) return;
    stage_tile<R, N, NT>(tile, u, leaf, links, metas);
    __syncthreads();
    double g = 0.0;
    for (int i = threadIdx.x; i < S::NI; i += NT) g = fmax(g, gradient_cell<R, N>(tile, i));
    g = block_max<NT>(g, red);
    if (threadIdx.x == 0) flags[leaf] = decide(g, key_level(keys[leaf]), rthr, cthr, min_level, max_level);
}

// Maximum-value criterion over the interior and the H-deep face halos,
// starting from 0.
template <int R, int N, int H, int NT>
__global__ __launch_bounds__(NT) void k_criterion_max(const double* __restrict__ u,
                                                      const uint64_t* __restrict__ keys,
                                                      const face_link* __restrict__ links,
                                                      const uint32_t* __restrict__ metas, double rthr, double cthr,
                                                      int min_level, int max_level, int8_t* __restrict__ flags,
                                                      uint32_t nleaves)
{
    constexpr int  P    = N + 2 * H;
    constexpr int  FLAT = (R == 2) ? P * P : P * P * P;
    __shared__ double red[NT];
    const uint32_t leaf = blockIdx.x;
    if (leaf >= nleaves) return;
    double m = 0.0;
    for (int i = threadIdx.x; i < FLAT; i += NT)
    {
        int          kind;
        const double v = padded_value<R, N, H>(u, leaf, links, metas, i, &kind);
        if (kind < 2) m = fmax(m, v);
    }
    m = block_max<NT>(m, red);
    if (threadIdx.x == 0) flags[leaf] = decide(m, key_level(keys[leaf]), rthr, cthr, min_level, max_level);
}

} // namespace amrhip

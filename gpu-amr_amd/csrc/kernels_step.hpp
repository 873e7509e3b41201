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
    double vel[3];
    double hs[3];
    double dtdx[kMaxLevels][3];
};

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

// Upwind update of interior cell i from the staged tile: the expression
// order of amr_solver::time_step_cpu (amr_solver.hpp:302-347) and
// AdvectionPhysics::rusanovFlux (AdvectionPhysics.hpp:45-66):
//   total = 0; total -= dt/dx_d * (F(u, u_R) - F(u_L, u)) for d = x, y(, z);
//   F(a, b) = 0.5 * (a v + b v) - (0.5 |v|) (b - a); out = u + total.
template <int R, int N>
__host__ __device__ inline double upwind_cell(const double* tile, int i, const double* dtdx, const double* vel,
                                              const double* hs)
{
    using T = tile_shape<R, N>;
    int c[R];
    patch_shape<R, N>::coords(i, c);
    const int    t     = T::idx(c);
    const double uc    = tile[t];
    double       total = 0.0;
    for (int d = 0; d < R; ++d)
    {
        const int    s  = T::stride(R - 1 - d); // physical axis d
        const double ul = tile[t - s];
        const double ur = tile[t + s];
        const double v  = vel[d];
        const double fL = 0.5 * (ul * v + uc * v) - hs[d] * (uc - ul);
        const double fR = 0.5 * (uc * v + ur * v) - hs[d] * (ur - uc);
        total -= dtdx[d] * (fR - fL);
    }
    return uc + total;
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
                                             const uint32_t* __restrict__ metas, step_params prm, uint32_t nleaves)
{
    using S = patch_shape<R, N>;
    __shared__ double tile[tile_shape<R, N>::TS];
    const uint32_t leaf = xcd_item(blockIdx.x, gridDim.x);
    if (leaf >= nleaves) return;
    stage_tile<R, N, NT>(tile, u, leaf, links, metas);
    __syncthreads();
    const int lv = key_level(keys[leaf]);
    double    dtdx[R], vel[R], hs[R];
    for (int d = 0; d < R; ++d)
    {
        dtdx[d] = prm.dtdx[lv][d];
        vel[d]  = prm.vel[d];
        hs[d]   = prm.hs[d];
    }
    double* out = un + static_cast<size_t>(leaf) * S::NI;
    for (int i = threadIdx.x; i < S::NI; i += NT) out[i] = upwind_cell<R, N>(tile, i, dtdx, vel, hs);
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

// kernels_euler.hpp -- compressible Euler equations through the same
// patch-AMR step: the Rusanov (local Lax-Friedrichs) flux of
// include/solver/EulerPhysics.hpp, the time step of amr_solver::time_step_cpu
// over R + 2 conservative fields, and the CFL time step reduced on the
// device.
//
// Fields are stored one after another (field-major): field v of leaf slot s
// starts at u + v * fstride + s * n^R, so the ghost rules, the regrid
// transfer and the criteria of the scalar path apply to each field as they
// are.  Every expression keeps the reference's order (no contraction), so
// results are bit-identical to the CPU oracle.
#pragma once

#include "kernels_step.hpp"

#pragma clang fp contract(off)

namespace amrhip
{

// Per-launch constants of the Euler kernels: gamma, the cell sizes of every
// level (physics_system::cell_sizes, formed on the host), the sweep direction.
struct euler_params
{
    double   gamma;
    double   cs[kMaxLevels][3];
    uint32_t reverse;
    uint32_t next_dt; // step kernels: fold the next step's dt minimum over their output
};

// Device scalars of one advance_batch_async: the running minimum of
// dx / speed (as the bits of a non-negative double, so that an integer
// atomicMin orders it), the dt of the current step, the summed dt, the
// remaining time, the executed step count and a stop flag (the CPU loop's
// break, amr_solver.hpp:225-241).
struct euler_scalars
{
    uint64_t dtmin_bits;
    double   step_dt;
    double   acc;
    double   remaining;
    uint64_t executed;
    uint32_t stopped;
    uint32_t pad;
};

template <int R>
struct euler_flux
{
    static constexpr int NV = R + 2;

    // primitiveToConservative (EulerPhysics.hpp:39-62).
    __host__ __device__ static void prim_to_cons(const double* prim, double* cons, double gamma)
    {
        const double rho = prim[0];
        const double p   = prim[R + 1];
        cons[0]          = rho;
        double vsq       = 0.0;
        for (int d = 0; d < R; ++d)
        {
            const double vel = prim[1 + d];
            cons[1 + d]      = rho * vel;
            vsq += vel * vel;
        }
        cons[R + 1] = p / (gamma - 1.0) + 0.5 * rho * vsq;
    }

    // 1/rho, pressure and sound speed of one conservative state, as
    // rusanovFlux and getMaxSpeed form them (EulerPhysics.hpp:80-94,
    // 140-152); the velocity along a direction is U[1 + dir] * inv_rho.  A
    // state is formed once per cell and reused for every face and direction
    // (the same operands give the same bits).
    struct state
    {
        double inv_rho, p, a;
    };
    __host__ __device__ static state state_of(const double* U, double gamma)
    {
        state        st;
        st.inv_rho = 1.0 / U[0];
        double kin = 0.0;
        for (int d = 0; d < R; ++d) kin += U[1 + d] * U[1 + d];
        kin *= 0.5 * st.inv_rho;
        st.p = (gamma - 1.0) * (U[R + 1] - kin);
        st.a = sqrt(gamma * st.p * st.inv_rho);
        return st;
    }

    // rusanovFlux (EulerPhysics.hpp:72-125) from the two states; smax =
    // std::max(l, r), i.e. (l < r) ? r : l.
    __host__ __device__ static void rusanov(const double* UL, state const& L, const double* UR, state const& Rs,
                                            double* flux, int dir)
    {
        const double ul = UL[1 + dir] * L.inv_rho, ur = UR[1 + dir] * Rs.inv_rho;
        const double sl = fabs(ul) + L.a, sr = fabs(ur) + Rs.a;
        const double smax = (sl < sr) ? sr : sl;
        flux[0]           = 0.5 * (UL[1 + dir] + UR[1 + dir] - smax * (UR[0] - UL[0]));
        for (int d = 0; d < R; ++d)
        {
            double fl = UL[1 + d] * ul;
            double fr = UR[1 + d] * ur;
            if (d == dir)
            {
                fl += L.p;
                fr += Rs.p;
            }
            flux[1 + d] = 0.5 * (fl + fr - smax * (UR[1 + d] - UL[1 + d]));
        }
        const double fle = ul * (UL[R + 1] + L.p);
        const double fre = ur * (UR[R + 1] + Rs.p);
        flux[R + 1]      = 0.5 * (fle + fre - smax * (UR[R + 1] - UL[R + 1]));
    }

    __host__ __device__ static void rusanov(const double* UL, const double* UR, double* flux, int dir, double gamma)
    {
        rusanov(UL, state_of(UL, gamma), UR, state_of(UR, gamma), flux, dir);
    }

    // getMaxSpeed (EulerPhysics.hpp:132-155).
    __host__ __device__ static double max_speed(const double* U, state const& st, int dir)
    {
        return fabs(U[1 + dir] * st.inv_rho) + st.a;
    }
};

// ghost_value for NV fields stored fstride apart: the face record is decoded
// once and the source cells of every field are loaded together.  Each
// field's arithmetic (the copy, the injection, the fine mean summed from 0.0
// in row-major order) is ghost_value's, so the values are the same bits.
template <int R, int NV>
__device__ inline void ghost_values(box<R> const& B, const double* __restrict__ u, size_t fstride,
                                    const face_link& L, uint32_t meta, int dir, const int* c, double* out)
{
    const uint32_t rel = rel_of(meta);
    if (rel == REL_SAME || rel == REL_COARSER)
    {
        const size_t at = ghost_source<R>(B, L.ids[0], meta, dir, c);
        for (int v = 0; v < NV; ++v) out[v] = u[v * fstride + at];
        return;
    }
    if (rel != REL_FINER)
    {
        for (int v = 0; v < NV; ++v) out[v] = 0.0;
        return;
    }
    const int dim = dir >> 1;
    const int nrm = (dir & 1) ? c[dim] - B.n(dim) : c[dim] + B.n(dim);
    int       k = 0, bit = 0, base[R];
    for (int d = 0; d < R; ++d)
    {
        const int w = (d == dim) ? nrm : c[d];
        base[d]     = (w * 2) % B.n(d);
        if (d != dim)
        {
            k |= (c[d] / B.half(d)) << bit;
            ++bit;
        }
    }
    const size_t q = static_cast<size_t>(L.ids[k]) * B.ni();
    double       sum[NV];
    for (int v = 0; v < NV; ++v) sum[v] = 0.0;
    for (int o = 0; o < box<R>::F; ++o)
    {
        int cc[R];
        for (int d = 0; d < R; ++d) cc[d] = base[d] + ((o >> (R - 1 - d)) & 1);
        const size_t at = q + B.lin(cc);
        for (int v = 0; v < NV; ++v) sum[v] += u[v * fstride + at];
    }
    for (int v = 0; v < NV; ++v) out[v] = sum[v] / static_cast<double>(box<R>::F);
}

template <int R, int N, int NV>
__device__ inline void ghost_values(const double* __restrict__ u, size_t fstride, const face_link& L, uint32_t meta,
                                    int dir, const int* c, double* out)
{
    constexpr box<R> B = box<R>::cube(N);
    ghost_values<R, NV>(B, u, fstride, L, meta, dir, c, out);
}

// The NV values of the cell with layout coordinates c of `leaf`: interior
// cells from the leaf, ghost cells (one coordinate just outside) through
// the ghost rules.
template <int R>
__device__ inline void euler_cell(box<R> const& B, const double* __restrict__ u, size_t fstride, uint32_t leaf,
                                  const face_link* __restrict__ links, const uint32_t* __restrict__ metas, const int* c,
                                  double* U)
{
    constexpr int NV = R + 2;
    int dir = -1;
    for (int d = 0; d < R; ++d)
    {
        if (c[d] < 0) dir = 2 * d;
        if (c[d] >= B.n(d)) dir = 2 * d + 1;
    }
    if (dir < 0)
    {
        const size_t at = static_cast<size_t>(leaf) * B.ni() + B.lin(c);
        for (int v = 0; v < NV; ++v) U[v] = u[v * fstride + at];
        return;
    }
    const size_t e = static_cast<size_t>(leaf) * box<R>::NDIR + dir;
    ghost_values<R, NV>(B, u, fstride, links[e], metas[e], dir, c, U);
}

// Every field's (n+2)^R LDS tile of `leaf` (field v at tile + v * TS): the interior cells, then the
// face-ghost cells with one decode of the face record per ghost cell for
// all fields.
template <int R, int N, int NT, int NV>
__device__ inline void stage_fields(double* tile, const double* __restrict__ u,
                                    size_t fstride, uint32_t leaf, const face_link* __restrict__ links,
                                    const uint32_t* __restrict__ metas)
{
    using S           = patch_shape<R, N>;
    using T           = tile_shape<R, N>;
    const size_t base = static_cast<size_t>(leaf) * S::NI;
    for (int i = threadIdx.x; i < S::NI; i += NT)
    {
        int c[R];
        S::coords(i, c);
        const int t = T::idx(c);
        for (int v = 0; v < NV; ++v) tile[v * T::TS + t] = u[v * fstride + base + i];
    }
    for (int g = threadIdx.x; g < S::NDIR * S::FACE; g += NT)
    {
        int          c[R];
        const int    dir = ghost_cell<R, N>(g, c);
        const size_t e   = static_cast<size_t>(leaf) * S::NDIR + dir;
        double       w[NV];
        ghost_values<R, N, NV>(u, fstride, links[e], metas[e], dir, c, w);
        const int t = T::idx(c);
        for (int v = 0; v < NV; ++v) tile[v * T::TS + t] = w[v];
    }
}

// The wave speed |u_d| + a of one cell along each dim (getMaxSpeed,
// EulerPhysics.hpp:132-155), folded into the running maxima smax.  A NaN
// speed never enters (the comparison fails), as the reference's
// `speed > 1e-12` test skips it.
template <int R>
__device__ inline void cell_speeds(const double* U, double gamma, double* smax)
{
    const auto st = euler_flux<R>::state_of(U, gamma);
    for (int d = 0; d < R; ++d)
    {
        const double speed = euler_flux<R>::max_speed(U, st, d);
        smax[d]            = (speed > smax[d]) ? speed : smax[d];
    }
}

// compute_time_step_cpu's minimum (amr_solver.hpp:380-400) over one
// workgroup's cells, which share a level and its cell sizes cs: the smallest
// cs[d] / S_d over the dims whose largest speed S_d exceeds 1e-12.
// Correctly rounded division is monotonic in the divisor, so cs[d] / S_d is
// the smallest of the cells' cs[d] / speed -- the same double as the
// per-cell minimum, for R divisions per workgroup instead of R per cell.
// Folded into the launch-wide minimum; every lane of the workgroup must call
// it.
template <int R, int NT>
__device__ inline void fold_speed_max(double* smax, const double* cs, euler_scalars* __restrict__ sc)
{
    for (int d = 0; d < R; ++d)
        for (int off = 32; off > 0; off >>= 1)
        {
            const double o = __shfl_down(smax[d], off, 64);
            smax[d]        = (o > smax[d]) ? o : smax[d];
        }
    // lane 0 of each wave holds the wave's maxima; a one-wave workgroup
    // needs no partials in LDS
    __shared__ double part[NT > 64 ? NT / 64 : 1][R]; // unused (and dropped) when NT == 64
    if constexpr (NT > 64)
    {
        if ((threadIdx.x & 63) == 0)
            for (int d = 0; d < R; ++d) part[threadIdx.x >> 6][d] = smax[d];
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        double dt = 1.7976931348623157e308;
        for (int d = 0; d < R; ++d)
        {
            double sd = smax[d];
            if constexpr (NT > 64)
            {
                sd = part[0][d];
                for (int w = 1; w < NT / 64; ++w) sd = (part[w][d] > sd) ? part[w][d] : sd;
            }
            if (sd > 1e-12)
            {
                const double q = cs[d] / sd;
                dt             = (q < dt) ? q : dt;
            }
        }
        // non-negative doubles order like their bit patterns; most
        // workgroups do not lower the minimum and skip the atomic (a stale
        // read only costs an extra one), which keeps thousands of
        // workgroups from queueing on one address
        const unsigned long long bits = static_cast<unsigned long long>(__double_as_longlong(dt));
        if (bits < *reinterpret_cast<volatile unsigned long long*>(&sc->dtmin_bits))
            atomicMin(reinterpret_cast<unsigned long long*>(&sc->dtmin_bits), bits);
    }
}

// compute_time_step_cpu (amr_solver.hpp:355-413): min over the leaf's cells
// and dims of dx_d / speed (speed > 1e-12), folded into the launch-wide
// minimum.  The minimum is exact and order-free, so any reduction order
// gives the oracle's value.
template <int R, int NT>
__device__ inline void euler_dt_leaf(box<R> const& B, const double* __restrict__ u, size_t fstride,
                                     const uint64_t* __restrict__ keys, euler_params const& prm,
                                     euler_scalars* __restrict__ sc, uint32_t leaf)
{
    constexpr int NV = R + 2;
    const int     ni = B.ni();
    const int     lv = key_level(keys[leaf]);
    double        smax[R];
    for (int d = 0; d < R; ++d) smax[d] = 0.0;
    for (int i = threadIdx.x; i < ni; i += NT)
    {
        double U[NV];
        for (int v = 0; v < NV; ++v) U[v] = u[v * fstride + static_cast<size_t>(leaf) * ni + i];
        cell_speeds<R>(U, prm.gamma, smax);
    }
    fold_speed_max<R, NT>(smax, prm.cs[lv], sc);
}

template <int R, int N, int NT>
__global__ __launch_bounds__(NT) void k_euler_dt(const double* __restrict__ u, size_t fstride,
                                                 const uint64_t* __restrict__ keys, euler_params prm,
                                                 euler_scalars* __restrict__ sc, uint32_t nleaves)
{
    constexpr box<R> B    = box<R>::cube(N);
    const uint32_t   leaf = xcd_item(blockIdx.x, gridDim.x);
    if (leaf >= nleaves) return;
    euler_dt_leaf<R, NT>(B, u, fstride, keys, prm, sc, leaf);
}

// k_euler_dt for patches of any shape.
template <int R, int NT>
__global__ __launch_bounds__(NT) void k_euler_dt_box(box<R> B, const double* __restrict__ u, size_t fstride,
                                                     const uint64_t* __restrict__ keys, euler_params prm,
                                                     euler_scalars* __restrict__ sc, uint32_t nleaves)
{
    const uint32_t leaf = xcd_item(blockIdx.x, gridDim.x);
    if (leaf >= nleaves) return;
    euler_dt_leaf<R, NT>(B, u, fstride, keys, prm, sc, leaf);
}

// The CPU loop's per-step dt logic (amr_solver.hpp:225-241) on the device:
// stop when the remaining time is used up, else dt = min(cfl * dtmin,
// remaining); reset the minimum for the next step.
__global__ void k_euler_finalize(euler_scalars* __restrict__ sc, double cfl)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double raw = __longlong_as_double(static_cast<long long>(sc->dtmin_bits));
    sc->dtmin_bits   = static_cast<uint64_t>(__double_as_longlong(1.7976931348623157e308));
    double step_dt   = 0.0;
    if (!sc->stopped && sc->remaining > 0.0)
    {
        const double dt = cfl * raw;
        step_dt         = (sc->remaining < dt) ? sc->remaining : dt;
    }
    if (step_dt <= 0.0)
    {
        sc->stopped = 1;
        sc->step_dt = 0.0;
        return;
    }
    sc->step_dt = step_dt;
    sc->acc += step_dt;
    sc->remaining -= step_dt;
    sc->executed += 1;
}

// Batch start: remaining time, zero sums, dt minimum reset (a kernel, not a
// memset, so that a captured graph replays it).
__global__ void k_euler_begin(euler_scalars* __restrict__ sc, double remaining)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    sc->dtmin_bits = static_cast<uint64_t>(__double_as_longlong(1.7976931348623157e308));
    sc->step_dt    = 0.0;
    sc->acc        = 0.0;
    sc->remaining  = remaining;
    sc->executed   = 0;
    sc->stopped    = 0;
}

// time_step_cpu (amr_solver.hpp:265-353) for the Euler fields: per cell and
// physical dim d, fL = F(U[c - e_d], U[c]), fR = F(U[c], U[c + e_d]),
// total[k] -= dt / dx_d * (fR[k] - fL[k]); out = U + total.  A step the dt
// logic stopped copies the state (the CPU loop runs no step).  Neighbours
// are read from memory.
template <int R, int NT>
__device__ inline void euler_step_leaf(box<R> const& B, const double* __restrict__ u, double* __restrict__ un,
                                       size_t fstride, const uint64_t* __restrict__ keys,
                                       const face_link* __restrict__ links, const uint32_t* __restrict__ metas,
                                       euler_params const& prm, euler_scalars* __restrict__ sc, uint32_t leaf)
{
    constexpr int NV   = R + 2;
    const int     ni   = B.ni();
    const double  dt   = sc->step_dt;
    const size_t  base = static_cast<size_t>(leaf) * ni;
    if (!(dt > 0.0))
    {
        for (int i = threadIdx.x; i < ni; i += NT)
            for (int v = 0; v < NV; ++v) un[v * fstride + base + i] = u[v * fstride + base + i];
        return;
    }
    const int lv = key_level(keys[leaf]);
    double    dtdx[R];
    for (int d = 0; d < R; ++d) dtdx[d] = dt / prm.cs[lv][d];
    double smax[R]; // next step's largest speeds over this workgroup's cells
    for (int d = 0; d < R; ++d) smax[d] = 0.0;
    for (int i = threadIdx.x; i < ni; i += NT)
    {
        int c[R];
        B.coords(i, c);
        double Uc[NV], total[NV];
        for (int v = 0; v < NV; ++v)
        {
            Uc[v]    = u[v * fstride + base + i];
            total[v] = 0.0;
        }
        const auto stc = euler_flux<R>::state_of(Uc, prm.gamma);
        for (int d = 0; d < R; ++d)
        {
            const int j = R - 1 - d; // layout dim of physical axis d
            double    UL[NV], UR[NV], fL[NV], fR[NV];
            c[j] -= 1;
            euler_cell<R>(B, u, fstride, leaf, links, metas, c, UL);
            c[j] += 2;
            euler_cell<R>(B, u, fstride, leaf, links, metas, c, UR);
            c[j] -= 1;
            euler_flux<R>::rusanov(UL, euler_flux<R>::state_of(UL, prm.gamma), Uc, stc, fL, d);
            euler_flux<R>::rusanov(Uc, stc, UR, euler_flux<R>::state_of(UR, prm.gamma), fR, d);
            for (int k = 0; k < NV; ++k) total[k] -= dtdx[d] * (fR[k] - fL[k]);
        }
        double o[NV];
        for (int v = 0; v < NV; ++v)
        {
            o[v]                       = Uc[v] + total[v];
            un[v * fstride + base + i] = o[v];
        }
        if (prm.next_dt) cell_speeds<R>(o, prm.gamma, smax);
    }
    if (prm.next_dt) fold_speed_max<R, NT>(smax, prm.cs[key_level(keys[leaf])], sc);
}

template <int R, int N, int NT>
__global__ __launch_bounds__(NT) void k_euler_step(const double* __restrict__ u, double* __restrict__ un,
                                                   size_t fstride, const uint64_t* __restrict__ keys,
                                                   const face_link* __restrict__ links,
                                                   const uint32_t* __restrict__ metas, euler_params prm,
                                                   euler_scalars* __restrict__ sc, uint32_t nleaves)
{
    constexpr box<R> B    = box<R>::cube(N);
    const uint32_t   leaf = sweep_item(blockIdx.x, gridDim.x, prm.reverse);
    if (leaf >= nleaves) return;
    euler_step_leaf<R, NT>(B, u, un, fstride, keys, links, metas, prm, sc, leaf);
}

// k_euler_step for patches of any shape.
template <int R, int NT>
__global__ __launch_bounds__(NT) void k_euler_step_box(box<R> B, const double* __restrict__ u, double* __restrict__ un,
                                                       size_t fstride, const uint64_t* __restrict__ keys,
                                                       const face_link* __restrict__ links,
                                                       const uint32_t* __restrict__ metas, euler_params prm,
                                                       euler_scalars* __restrict__ sc, uint32_t nleaves)
{
    const uint32_t leaf = sweep_item(blockIdx.x, gridDim.x, prm.reverse);
    if (leaf >= nleaves) return;
    euler_step_leaf<R, NT>(B, u, un, fstride, keys, links, metas, prm, sc, leaf);
}

// k_euler_step with the leaf's fields staged in LDS: every field's (n+2)^R
// tile (interior plus face ghosts, gathered once per ghost cell by the same
// rules) is loaded cooperatively, then each cell reads its neighbours from
// LDS.  For patch shapes whose tiles fit (3D 8^3: 40 KB).
template <int R, int N, int NT>
__global__ __launch_bounds__(NT) void k_euler_step_tile(const double* __restrict__ u, double* __restrict__ un,
                                                        size_t fstride, const uint64_t* __restrict__ keys,
                                                        const face_link* __restrict__ links,
                                                        const uint32_t* __restrict__ metas, euler_params prm,
                                                        euler_scalars* __restrict__ sc, uint32_t nleaves)
{
    using S          = patch_shape<R, N>;
    using T          = tile_shape<R, N>;
    constexpr int NV = R + 2;
    __shared__ double tile[NV][T::TS];
    const uint32_t leaf = sweep_item(blockIdx.x, gridDim.x, prm.reverse);
    if (leaf >= nleaves) return;
    const double dt   = sc->step_dt;
    const size_t base = static_cast<size_t>(leaf) * S::NI;
    if (!(dt > 0.0))
    {
        for (int i = threadIdx.x; i < S::NI; i += NT)
            for (int v = 0; v < NV; ++v) un[v * fstride + base + i] = u[v * fstride + base + i];
        return;
    }
    stage_fields<R, N, NT, NV>(&tile[0][0], u, fstride, leaf, links, metas);
    __syncthreads();
    const int lv = key_level(keys[leaf]);
    double    dtdx[R];
    for (int d = 0; d < R; ++d) dtdx[d] = dt / prm.cs[lv][d];
    double smax[R]; // next step's largest speeds over this workgroup's cells
    for (int d = 0; d < R; ++d) smax[d] = 0.0;
    for (int i = threadIdx.x; i < S::NI; i += NT)
    {
        int c[R];
        S::coords(i, c);
        const int t = T::idx(c);
        double    Uc[NV], total[NV];
        for (int v = 0; v < NV; ++v)
        {
            Uc[v]    = tile[v][t];
            total[v] = 0.0;
        }
        const auto stc = euler_flux<R>::state_of(Uc, prm.gamma);
        for (int d = 0; d < R; ++d)
        {
            const int s = T::stride(R - 1 - d); // physical axis d
            double    UL[NV], UR[NV], fL[NV], fR[NV];
            for (int v = 0; v < NV; ++v)
            {
                UL[v] = tile[v][t - s];
                UR[v] = tile[v][t + s];
            }
            euler_flux<R>::rusanov(UL, euler_flux<R>::state_of(UL, prm.gamma), Uc, stc, fL, d);
            euler_flux<R>::rusanov(Uc, stc, UR, euler_flux<R>::state_of(UR, prm.gamma), fR, d);
            for (int k = 0; k < NV; ++k) total[k] -= dtdx[d] * (fR[k] - fL[k]);
        }
        double o[NV];
        for (int v = 0; v < NV; ++v)
        {
            o[v]                              = Uc[v] + total[v];
            un[v * fstride + base + i] = o[v];
        }
        if (prm.next_dt) cell_speeds<R>(o, prm.gamma, smax);
    }
    if (prm.next_dt) fold_speed_max<R, NT>(smax, prm.cs[key_level(keys[leaf])], sc);
}

// A value every lane holds alike, read from the first active lane into
// scalar registers (the VALU result of a uniform division otherwise holds
// vector registers).
__device__ inline double wave_uniform(double v)
{
    const uint64_t b  = __double_as_longlong(v);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(b));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(b >> 32));
    return __longlong_as_double(static_cast<long long>((static_cast<uint64_t>(hi) << 32) | lo));
}

// The five conservative values and the state of one cell moved across
// lanes as 64-bit DPP rows (lane_plus1) or LDS permutes (__shfl_down by 8),
// every lane active.
struct euler_cell3
{
    double U[5];
    double inv_rho, p, a;
};

__device__ inline euler_cell3 cell_lane_plus1(euler_cell3 const& c)
{
    euler_cell3 o;
#pragma unroll
    for (int v = 0; v < 5; ++v) o.U[v] = lane_plus1(c.U[v]);
    o.inv_rho = lane_plus1(c.inv_rho);
    o.p       = lane_plus1(c.p);
    o.a       = lane_plus1(c.a);
    return o;
}

__device__ inline euler_cell3 cell_lane_down8(euler_cell3 const& c)
{
    euler_cell3 o;
#pragma unroll
    for (int v = 0; v < 5; ++v) o.U[v] = __shfl_down(c.U[v], 8, 64);
    o.inv_rho = __shfl_down(c.inv_rho, 8, 64);
    o.p       = __shfl_down(c.p, 8, 64);
    o.a       = __shfl_down(c.a, 8, 64);
    return o;
}

// 3D 8^3 patches, one wave per leaf, lane l = 8 y + x owning the z-column
// (y, x), walked z = 0..7 with the column's cells in registers: each cell's
// state (1/rho, p, a) is formed once, and each face's Rusanov flux is
// evaluated once -- by the cell on its low side, from that cell's and its +x
// (next lane), +y (lane + 8) or +z (next plane) neighbour's values and
// states -- and handed to the cell on its high side (the previous lane, lane
// - 8, the next plane).  The fluxes through the patch's x and y faces are
// evaluated in a face-layout pass first (lane = 8 z + q: the ghost cell by
// the face rules, the boundary cell, their states, the flux) and read from
// LDS by the boundary lanes; the z faces' ghosts are the lane's own.  Every
// flux and state is formed from the same operands by the same expressions
// as in k_euler_step_tile, so the results are the same bits, for about half
// of its VALU instructions (the tile kernel forms each cell's state seven
// times and each face's flux twice).
// The face records are read once; unless a face has finer neighbours every
// ghost is a single cell and all ghost and boundary-cell loads are issued
// back to back, the +z ghost in the prefetch slot of plane 8.  Four waves
// per SIMD: at most 128 VGPRs (z loop unrolled, so the prefetch needs no
// register copies; the step's dt / dx in scalar registers) and 10 KiB of
// LDS (the boundary fluxes only).
template <bool NTS>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_euler_step_3d8(
    const double* __restrict__ u, double* __restrict__ un, size_t fstride, const uint64_t* __restrict__ keys,
    const face_link* __restrict__ links, const uint32_t* __restrict__ metas, euler_params prm,
    euler_scalars* __restrict__ sc, uint32_t nleaves)
{
    using F            = euler_flux<3>;
    constexpr int NV   = 5;
    constexpr int NI   = 512;
    __shared__ double fb[4][NV][64]; // boundary fluxes -y, +y (index z*8+x), -x, +x (index z*8+y)
    const uint32_t    leaf = sweep_item(blockIdx.x, gridDim.x, prm.reverse);
    if (leaf >= nleaves) return;
    const int    lane = threadIdx.x, y = lane >> 3, x = lane & 7;
    const double dt   = sc->step_dt;
    const size_t base = static_cast<size_t>(leaf) * NI;
    if (!(dt > 0.0))
    {
        for (int i = lane; i < NI; i += 64)
            for (int v = 0; v < NV; ++v) un[v * fstride + base + i] = u[v * fstride + base + i];
        return;
    }
    const face_link* L     = links + static_cast<size_t>(leaf) * 6;
    const uint32_t*  M     = metas + static_cast<size_t>(leaf) * 6;
    const double     gamma = prm.gamma;
    const int        p = lane >> 3, q = lane & 7;

    // The leaf's six face records, read once.
    uint32_t m[6];
    int32_t  nb[6];
#pragma unroll
    for (int k = 0; k < 6; ++k)
    {
        m[k]  = M[k];
        nb[k] = L[k].ids[0];
    }

    // ---- ghost cells and the boundary fluxes of the x and y faces ----------
    // Ghost k (direction k) of the lane: k = 0, 1 the -z / +z ghost of the
    // column (y, x) = (p, q); k = 2..5 the ghost of face cell (z, x) = (p, q)
    // of the -y / +y face, (z, y) = (p, q) of the -x / +x face.  bc[f] is
    // the patch cell on the inner side of side face f = k - 2.
    const int gc[6][3] = { { -1, p, q }, { 8, p, q }, { p, -1, q }, { p, 8, q }, { p, q, -1 }, { p, q, 8 } };
    const int bc[4]    = { p * 64 + q, p * 64 + 56 + q, p * 64 + q * 8, p * 64 + q * 8 + 7 };
    double    gzm[NV]; // the -z ghost (the +z ghost is loaded with the last planes)
    const bool simple = single_cell_ghosts<3>(m);
    auto      side_flux = [&](int f, const double* G, const double* C) {
        const auto sg = F::state_of(G, gamma), sb = F::state_of(C, gamma);
        double     fl[NV];
        if (f & 1)
            F::rusanov(C, sb, G, sg, fl, f < 2 ? 1 : 0);
        else
            F::rusanov(G, sg, C, sb, fl, f < 2 ? 1 : 0);
#pragma unroll
        for (int v = 0; v < NV; ++v) fb[f][v][lane] = fl[v];
    };
    if (simple)
    {
        // every ghost is one cell (same-level copy or coarser injection):
        // all 45 loads issued back to back
        double G[6][NV], C[4][NV];
#pragma unroll
        for (int k = 0; k < 6; ++k)
        {
            if (k == 1) continue;
            const size_t src = ghost_source<3, 8>(nb[k], m[k], k, gc[k]);
#pragma unroll
            for (int v = 0; v < NV; ++v) G[k][v] = u[v * fstride + src];
        }
#pragma unroll
        for (int f = 0; f < 4; ++f)
#pragma unroll
            for (int v = 0; v < NV; ++v) C[f][v] = u[v * fstride + base + bc[f]];
#pragma unroll
        for (int v = 0; v < NV; ++v) gzm[v] = G[0][v];
#pragma unroll
        for (int f = 0; f < 4; ++f) side_flux(f, G[2 + f], C[f]);
    }
    else
    {
        double G[NV];
        ghost_values<3, 8, NV>(u, fstride, L[0], m[0], 0, gc[0], gzm);
#pragma unroll
        for (int f = 0; f < 4; ++f)
        {
            double C[NV];
            ghost_values<3, 8, NV>(u, fstride, L[2 + f], m[2 + f], 2 + f, gc[2 + f], G);
#pragma unroll
            for (int v = 0; v < NV; ++v) C[v] = u[v * fstride + base + bc[f]];
            side_flux(f, G, C);
        }
    }

    // ---- the column, z = 0..7 ----------------------------------------------
    // each field's cells of the leaf (uniform pointers, the lane's cell
    // of plane z at [z * 64 + lane])
    const double* uf[NV];
    double*       of[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v)
    {
        uf[v] = u + v * fstride + base;
        of[v] = un + v * fstride + base;
    }
    euler_cell3 cur, nxt;
    double      pre[NV]; // the plane after the next one, loaded an iteration ahead
#pragma unroll
    for (int v = 0; v < NV; ++v)
    {
        cur.U[v] = uf[v][lane];
        pre[v]   = uf[v][64 + lane];
    }
    const int lv      = key_level(keys[leaf]);
    double    dtdx[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) dtdx[d] = wave_uniform(dt / prm.cs[lv][d]); // kept in scalar registers
    double    smax[3] = { 0.0, 0.0, 0.0 };
    {
        const auto st = F::state_of(cur.U, gamma);
        cur.inv_rho = st.inv_rho, cur.p = st.p, cur.a = st.a;
        const auto sg = F::state_of(gzm, gamma);
        double     fz[NV];
        F::rusanov(gzm, sg, cur.U, st, fz, 2);
        __syncthreads(); // the boundary fluxes
#pragma unroll
        for (int z = 0; z < 8; ++z)
        {
            const typename F::state sc_{ cur.inv_rho, cur.p, cur.a };
            // total[k] -= dt/dx_d (fR[k] - fL[k]) for d = x, y, z in that
            // order, each axis's fluxes dead before the next axis's are formed
            double total[NV];
#pragma unroll
            for (int k = 0; k < NV; ++k) total[k] = 0.0;
            {
                // x: the +x face from this lane and the next, the -x face from the previous lane
                const euler_cell3 xc = cell_lane_plus1(cur);
                double            fRx[NV];
                F::rusanov(cur.U, sc_, xc.U, typename F::state{ xc.inv_rho, xc.p, xc.a }, fRx, 0);
#pragma unroll
                for (int v = 0; v < NV; ++v) fRx[v] = (x == 7) ? fb[3][v][z * 8 + y] : fRx[v];
#pragma unroll
                for (int v = 0; v < NV; ++v)
                {
                    const double m1  = lane_minus1(fRx[v]);
                    const double fLx = (x == 0) ? fb[2][v][z * 8 + y] : m1;
                    total[v] -= dtdx[0] * (fRx[v] - fLx);
                }
            }
            {
                // y: the +y face from this lane and lane + 8, the -y face from lane - 8
                const euler_cell3 yc = cell_lane_down8(cur);
                double            fRy[NV];
                F::rusanov(cur.U, sc_, yc.U, typename F::state{ yc.inv_rho, yc.p, yc.a }, fRy, 1);
#pragma unroll
                for (int v = 0; v < NV; ++v) fRy[v] = (y == 7) ? fb[1][v][z * 8 + x] : fRy[v];
#pragma unroll
                for (int v = 0; v < NV; ++v)
                {
                    const double m8  = __shfl_up(fRy[v], 8, 64);
                    const double fLy = (y == 0) ? fb[0][v][z * 8 + x] : m8;
                    total[v] -= dtdx[1] * (fRy[v] - fLy);
                }
            }
            // z: the +z neighbour is the next plane (prefetched) or the +z ghost
            if (z < 7)
            {
#pragma unroll
                for (int v = 0; v < NV; ++v) nxt.U[v] = pre[v];
                if (z < 6)
                {
#pragma unroll
                    for (int v = 0; v < NV; ++v) pre[v] = uf[v][(z + 2) * 64 + lane];
                }
                else if (simple)
                {
                    // plane 8: the +z ghost, a single cell
                    const size_t src = ghost_source<3, 8>(nb[1], m[1], 1, gc[1]);
#pragma unroll
                    for (int v = 0; v < NV; ++v) pre[v] = u[v * fstride + src];
                }
            }
            else if (simple)
            {
#pragma unroll
                for (int v = 0; v < NV; ++v) nxt.U[v] = pre[v];
            }
            else
                ghost_values<3, 8, NV>(u, fstride, L[1], m[1], 1, gc[1], nxt.U);
            {
                const auto st = F::state_of(nxt.U, gamma);
                nxt.inv_rho = st.inv_rho, nxt.p = st.p, nxt.a = st.a;
            }
            double fzr[NV];
            F::rusanov(cur.U, sc_, nxt.U, typename F::state{ nxt.inv_rho, nxt.p, nxt.a }, fzr, 2);
            double o[NV];
#pragma unroll
            for (int k = 0; k < NV; ++k)
            {
                total[k] -= dtdx[2] * (fzr[k] - fz[k]);
                o[k] = cur.U[k] + total[k];
                store_cell<NTS>(of[k] + z * 64 + lane, o[k]);
            }
            if (prm.next_dt) cell_speeds<3>(o, gamma, smax);
#pragma unroll
            for (int v = 0; v < NV; ++v) fz[v] = fzr[v];
            cur = nxt;
        }
    }
    if (prm.next_dt) fold_speed_max<3, 64>(smax, prm.cs[lv], sc);
}

// The value of lane l + 1 (wave_shl) or l - 1 (wave_shr) across the whole
// wave; the lane whose source lies outside the wave (63, resp. 0) takes
// `fill`.  Every lane must be active.
__device__ inline double wave_plus1_fill(double v, double fill)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, v), f = __builtin_bit_cast(uint64_t, fill);
    const int      l = __builtin_amdgcn_update_dpp(static_cast<int>(f), static_cast<int>(b), 0x130, 0xF, 0xF, false);
    const int      h =
        __builtin_amdgcn_update_dpp(static_cast<int>(f >> 32), static_cast<int>(b >> 32), 0x130, 0xF, 0xF, false);
    return __builtin_bit_cast(double, (static_cast<uint64_t>(static_cast<uint32_t>(h)) << 32) | static_cast<uint32_t>(l));
}

__device__ inline double wave_minus1_fill(double v, double fill)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, v), f = __builtin_bit_cast(uint64_t, fill);
    const int      l = __builtin_amdgcn_update_dpp(static_cast<int>(f), static_cast<int>(b), 0x138, 0xF, 0xF, false);
    const int      h =
        __builtin_amdgcn_update_dpp(static_cast<int>(f >> 32), static_cast<int>(b >> 32), 0x138, 0xF, 0xF, false);
    return __builtin_bit_cast(double, (static_cast<uint64_t>(static_cast<uint32_t>(h)) << 32) | static_cast<uint32_t>(l));
}

// The four values and the state of one 2D cell.
struct euler_cell2
{
    double U[4];
    double inv_rho, p, a;
};

// 2D 64^2 patches (the reference's 2D benchmark shape): one wave per strip
// of B rows, lane = x, the strip's rows walked y = r0 .. r0 + B - 1 with the
// row's cells in registers.  Each cell's state is formed once and each face
// flux is evaluated once: the +x flux by the cell on the face's low side from
// its own values and state and those of the next lane (a wave-wide DPP
// shift; lane 63 takes the +x ghost cell as the shift's fill), handed to the
// next lane as its -x flux (lane 0 takes the -x boundary flux as the fill);
// the +y flux of a row is the -y flux of the next row.  The -x boundary
// fluxes and the +x ghost cells of the strip's rows are formed first, one
// row per lane.  The fluxes and states come from the same operands by the
// same expressions as in k_euler_step_tile (the update of time_step_cpu,
// amr_solver.hpp:265-353, with rusanovFlux, EulerPhysics.hpp:72-125), so the
// results are the same bits.  Five waves per SIMD (at most 96 VGPRs: rows
// unrolled, uniform field pointers, dt / dx in scalar registers).
template <int B>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5))) void k_euler_step_2d64(
    const double* __restrict__ u, double* __restrict__ un, size_t fstride, const uint64_t* __restrict__ keys,
    const face_link* __restrict__ links, const uint32_t* __restrict__ metas, euler_params prm,
    euler_scalars* __restrict__ sc, uint32_t nleaves)
{
    static_assert(B >= 2 && 64 % B == 0 && 2 * B <= 64, "strips of at least two rows");
    using F           = euler_flux<2>;
    constexpr int NV  = 4;
    constexpr int N   = 64;
    constexpr int NI  = N * N;
    constexpr int NS  = N / B;
    __shared__ double xg[NV + 3][B]; // +x ghost cell of each strip row: values, 1/rho, p, a
    __shared__ double xf[NV][B];     // -x boundary flux of each strip row
    const uint32_t    item = sweep_item(blockIdx.x, gridDim.x, prm.reverse);
    const uint32_t    leaf = item / NS;
    if (leaf >= nleaves) return;
    const int    r0   = static_cast<int>(item % NS) * B;
    const int    x    = threadIdx.x;
    const double dt   = sc->step_dt;
    const size_t base = static_cast<size_t>(leaf) * NI;
    if (!(dt > 0.0))
    {
        for (int k = 0; k < B; ++k)
            for (int v = 0; v < NV; ++v)
                un[v * fstride + base + (r0 + k) * N + x] = u[v * fstride + base + (r0 + k) * N + x];
        return;
    }
    const face_link* L     = links + static_cast<size_t>(leaf) * 4;
    const uint32_t*  M     = metas + static_cast<size_t>(leaf) * 4;
    const double     gamma = prm.gamma;
    // each field's cells of the leaf (uniform pointers; the lane's cell of
    // row y at [y * N + x])
    const double* uf[NV];
    double*       of[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v)
    {
        uf[v] = u + v * fstride + base;
        of[v] = un + v * fstride + base;
    }
    euler_cell2 cur, nxt;
    double      pre[NV];   // the row after the next one, loaded an iteration ahead
    double      below[NV]; // the row under the strip
#pragma unroll
    for (int v = 0; v < NV; ++v)
    {
        cur.U[v] = uf[v][r0 * N + x];
        pre[v]   = uf[v][(r0 + 1) * N + x];
    }
    if (r0 > 0)
    {
#pragma unroll
        for (int v = 0; v < NV; ++v) below[v] = uf[v][(r0 - 1) * N + x];
    }
    else
    {
        const int c[2] = { -1, x };
        ghost_values<2, N, NV>(u, fstride, L[0], M[0], 0, c, below);
    }

    // ---- x boundary, one strip row per lane: lanes 0 .. B-1 the -x flux,
    // lanes B .. 2B-1 the +x ghost cell and its state
    if (x < 2 * B)
    {
        const int side = x >= B ? 1 : 0, k = x - side * B, row = r0 + k, dir = 2 + side;
        const int c[2] = { row, side ? N : -1 };
        double    G[NV];
        ghost_values<2, N, NV>(u, fstride, L[dir], M[dir], dir, c, G);
        const auto sg = F::state_of(G, gamma);
        if (side)
        {
#pragma unroll
            for (int v = 0; v < NV; ++v) xg[v][k] = G[v];
            xg[NV][k] = sg.inv_rho, xg[NV + 1][k] = sg.p, xg[NV + 2][k] = sg.a;
        }
        else
        {
            double       C[NV], fl[NV];
            const size_t at = base + static_cast<size_t>(row) * N;
#pragma unroll
            for (int v = 0; v < NV; ++v) C[v] = u[v * fstride + at];
            F::rusanov(G, sg, C, F::state_of(C, gamma), fl, 0);
#pragma unroll
            for (int v = 0; v < NV; ++v) xf[v][k] = fl[v];
        }
    }

    // ---- the strip's rows ---------------------------------------------------
    const int    lv      = key_level(keys[leaf]);
    const double dtdx[2] = { wave_uniform(dt / prm.cs[lv][0]), wave_uniform(dt / prm.cs[lv][1]) };
    double       smax[2] = { 0.0, 0.0 };
    double       fy[NV]; // the -y flux of the current row
    {
        const auto st = F::state_of(cur.U, gamma);
        cur.inv_rho = st.inv_rho, cur.p = st.p, cur.a = st.a;
        F::rusanov(below, F::state_of(below, gamma), cur.U, st, fy, 1);
    }
    __syncthreads(); // the x boundary values
#pragma unroll
    for (int k = 0; k < B; ++k)
    {
        const int               y = r0 + k;
        const typename F::state sc_{ cur.inv_rho, cur.p, cur.a };
        // total[v] -= dt/dx_d (fR[v] - fL[v]) for d = x, then y
        double total[NV];
        {
            euler_cell2 xc;
#pragma unroll
            for (int v = 0; v < NV; ++v) xc.U[v] = wave_plus1_fill(cur.U[v], xg[v][k]);
            xc.inv_rho = wave_plus1_fill(cur.inv_rho, xg[NV][k]);
            xc.p       = wave_plus1_fill(cur.p, xg[NV + 1][k]);
            xc.a       = wave_plus1_fill(cur.a, xg[NV + 2][k]);
            double fRx[NV];
            F::rusanov(cur.U, sc_, xc.U, typename F::state{ xc.inv_rho, xc.p, xc.a }, fRx, 0);
#pragma unroll
            for (int v = 0; v < NV; ++v)
            {
                const double fLx = wave_minus1_fill(fRx[v], xf[v][k]);
                total[v]         = 0.0;
                total[v] -= dtdx[0] * (fRx[v] - fLx);
            }
        }
        // the next row: prefetched; the row after it is the patch's or, past
        // the patch's last row, the +y ghost row
#pragma unroll
        for (int v = 0; v < NV; ++v) nxt.U[v] = pre[v];
        if (k + 2 <= B)
        {
            if (y + 2 < N)
            {
#pragma unroll
                for (int v = 0; v < NV; ++v) pre[v] = uf[v][(y + 2) * N + x];
            }
            else
            {
                const int c[2] = { N, x };
                ghost_values<2, N, NV>(u, fstride, L[1], M[1], 1, c, pre);
            }
        }
        {
            const auto st = F::state_of(nxt.U, gamma);
            nxt.inv_rho = st.inv_rho, nxt.p = st.p, nxt.a = st.a;
        }
        double fyr[NV];
        F::rusanov(cur.U, sc_, nxt.U, typename F::state{ nxt.inv_rho, nxt.p, nxt.a }, fyr, 1);
        double o[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v)
        {
            total[v] -= dtdx[1] * (fyr[v] - fy[v]);
            o[v] = cur.U[v] + total[v];
            of[v][y * N + x] = o[v];
        }
        if (prm.next_dt) cell_speeds<2>(o, gamma, smax);
#pragma unroll
        for (int v = 0; v < NV; ++v) fy[v] = fyr[v];
        cur = nxt;
    }
    if (prm.next_dt) fold_speed_max<2, 64>(smax, prm.cs[lv], sc);
}

// Strip of B layers (along layout dim 0, the slowest) of an n^R patch with
// one ghost layer all round: (B + 2) * (n + 2)^(R-1) LDS cells per field.
template <int R, int N, int B>
struct strip_shape
{
    static_assert(N % B == 0, "strips tile the patch");
    static constexpr int TP = N + 2;
    static constexpr int PL = (R == 2) ? TP : TP * TP; // tile cells per layer
    static constexpr int CL = (R == 2) ? N : N * N;    // patch cells per layer
    static constexpr int TS = (B + 2) * PL;
    static constexpr int NS = N / B; // strips per patch
    // tile index of patch cell c in the strip whose first layer is r0
    __device__ static int idx(const int* c, int r0)
    {
        int l = c[0] - r0 + 1;
        for (int d = 1; d < R; ++d) l = l * TP + (c[d] + 1);
        return l;
    }
    // stride of layout dim j inside the tile
    __host__ __device__ static constexpr int stride(int j) { return j == 0 ? PL : (j == R - 1 ? 1 : TP); }
};

// Cell q of a strip tile (strip_shape<R, N, B>, first layer r0): its patch
// coordinates in c; returns -1 for a cell of the patch, the face direction
// for a face-ghost cell and -2 for an edge or corner cell (never read).
template <int R, int N, int B>
__device__ inline int strip_cell(int q, int r0, int* c)
{
    using T = strip_shape<R, N, B>;
    int r   = q;
    for (int d = R - 1; d >= 0; --d)
    {
        const int ext = (d == 0) ? B + 2 : T::TP;
        c[d]          = r % ext - 1;
        r /= ext;
    }
    c[0] += r0;
    int outside = 0, dir = 0;
    for (int d = 0; d < R; ++d)
    {
        if (c[d] < 0)
        {
            ++outside;
            dir = 2 * d;
        }
        else if (c[d] >= N)
        {
            ++outside;
            dir = 2 * d + 1;
        }
    }
    return outside == 0 ? -1 : (outside == 1 ? dir : -2);
}

// The update of one staged strip (tile and LDS states formed): every cell's
// Rusanov update as in k_euler_step_tile, stored to the next state, and the
// new values' wave speeds folded into smax (the next step's dt).
template <int R, int N, int B, int NT>
__device__ inline void strip_update(const double (*tile)[strip_shape<R, N, B>::TS], const double* st_inv,
                                    const double* st_p, const double* st_a, int r0, size_t first, int lv, double dt,
                                    euler_params const& prm, double* __restrict__ un, size_t fstride, double* smax)
{
    using T          = strip_shape<R, N, B>;
    using F          = euler_flux<R>;
    constexpr int NV = R + 2;
    double        dtdx[R];
    for (int d = 0; d < R; ++d) dtdx[d] = dt / prm.cs[lv][d];
    for (int i = threadIdx.x; i < B * T::CL; i += NT)
    {
        int c[R], r = i % T::CL;
        c[0] = r0 + i / T::CL;
        for (int d = R - 1; d >= 1; --d)
        {
            c[d] = r % N;
            r /= N;
        }
        const int t = T::idx(c, r0);
        double    Uc[NV], total[NV];
        for (int v = 0; v < NV; ++v)
        {
            Uc[v]    = tile[v][t];
            total[v] = 0.0;
        }
        const typename F::state stc{st_inv[t], st_p[t], st_a[t]};
        for (int d = 0; d < R; ++d)
        {
            const int s = T::stride(R - 1 - d); // physical axis d
            double    UL[NV], UR[NV], fL[NV], fR[NV];
            for (int v = 0; v < NV; ++v)
            {
                UL[v] = tile[v][t - s];
                UR[v] = tile[v][t + s];
            }
            const typename F::state stl{st_inv[t - s], st_p[t - s], st_a[t - s]};
            const typename F::state str{st_inv[t + s], st_p[t + s], st_a[t + s]};
            F::rusanov(UL, stl, Uc, stc, fL, d);
            F::rusanov(Uc, stc, UR, str, fR, d);
            for (int k = 0; k < NV; ++k) total[k] -= dtdx[d] * (fR[k] - fL[k]);
        }
        double o[NV];
        for (int v = 0; v < NV; ++v)
        {
            o[v]                     = Uc[v] + total[v];
            un[v * fstride + first + i] = o[v];
        }
        if (prm.next_dt) cell_speeds<R>(o, prm.gamma, smax);
    }
}

// Every tile cell's state (1/rho, pressure, sound speed) into LDS (edge /
// corner cells hold stale values; their states are never read).
template <int R, int N, int B, int NT>
__device__ inline void strip_states(const double (*tile)[strip_shape<R, N, B>::TS], double* st_inv, double* st_p,
                                    double* st_a, double gamma)
{
    constexpr int NV = R + 2;
    for (int q = threadIdx.x; q < strip_shape<R, N, B>::TS; q += NT)
    {
        double U[NV];
        for (int v = 0; v < NV; ++v) U[v] = tile[v][q];
        const auto st = euler_flux<R>::state_of(U, gamma);
        st_inv[q]     = st.inv_rho;
        st_p[q]       = st.p;
        st_a[q]       = st.a;
    }
}

// k_euler_step_tile for patches whose whole tile does not fit in LDS (2D
// 64^2: 139 KB): one workgroup per strip of B layers, so that a leaf is n / B
// workgroups.  The strip tile holds the strip's cells, the neighbouring
// layers of the same patch (or the face ghosts beyond the patch) and the
// face ghosts along the other dims.  The state (1/rho, pressure, sound
// speed) of every tile cell is formed once into LDS and read by the cell and
// its neighbours; every update reads the same operands as in
// k_euler_step_tile, so the results are the same bits.
template <int R, int N, int B, int NT>
__global__ __launch_bounds__(NT) void k_euler_step_strip(const double* __restrict__ u, double* __restrict__ un,
                                                         size_t fstride, const uint64_t* __restrict__ keys,
                                                         const face_link* __restrict__ links,
                                                         const uint32_t* __restrict__ metas, euler_params prm,
                                                         euler_scalars* __restrict__ sc, uint32_t nleaves)
{
    using S          = patch_shape<R, N>;
    using T          = strip_shape<R, N, B>;
    constexpr int NV = R + 2;
    __shared__ double tile[NV][T::TS];
    __shared__ double st_inv[T::TS], st_p[T::TS], st_a[T::TS];
    const uint32_t item = sweep_item(blockIdx.x, gridDim.x, prm.reverse);
    const uint32_t leaf = item / T::NS;
    if (leaf >= nleaves) return;
    const int    r0    = static_cast<int>(item % T::NS) * B;
    const double dt    = sc->step_dt;
    const size_t base  = static_cast<size_t>(leaf) * S::NI;
    const size_t first = base + static_cast<size_t>(r0) * T::CL; // the strip's first cell
    if (!(dt > 0.0))
    {
        for (int i = threadIdx.x; i < B * T::CL; i += NT)
            for (int v = 0; v < NV; ++v) un[v * fstride + first + i] = u[v * fstride + first + i];
        return;
    }
    for (int q = threadIdx.x; q < T::TS; q += NT)
    {
        int       c[R];
        const int kind = strip_cell<R, N, B>(q, r0, c);
        if (kind == -2) continue; // edge / corner cells are never read
        double w[NV];
        if (kind == -1)
        {
            const size_t at = base + S::lin(c);
            for (int v = 0; v < NV; ++v) w[v] = u[v * fstride + at];
        }
        else
        {
            const size_t e = static_cast<size_t>(leaf) * S::NDIR + kind;
            ghost_values<R, N, NV>(u, fstride, links[e], metas[e], kind, c, w);
        }
        for (int v = 0; v < NV; ++v) tile[v][q] = w[v];
    }
    __syncthreads();
    strip_states<R, N, B, NT>(tile, st_inv, st_p, st_a, prm.gamma);
    __syncthreads();
    double smax[R]; // next step's largest speeds over this workgroup's cells
    for (int d = 0; d < R; ++d) smax[d] = 0.0;
    strip_update<R, N, B, NT>(tile, st_inv, st_p, st_a, r0, first, key_level(keys[leaf]), dt, prm, un, fstride, smax);
    if (prm.next_dt) fold_speed_max<R, NT>(smax, prm.cs[key_level(keys[leaf])], sc);
}

// Largest strip depth B (a divisor of n) whose strip tiles of the R + 2
// fields and the three state arrays fit in `bytes` of LDS; 0 if none does.
template <int R, int N>
constexpr int euler_strip_depth(size_t bytes)
{
    for (int b = N; b >= 1; --b)
        if (N % b == 0 && (R + 5) * (b + 2) * static_cast<size_t>(R == 2 ? N + 2 : (N + 2) * (N + 2)) * 8 <= bytes)
            return b;
    return 0;
}

} // namespace amrhip

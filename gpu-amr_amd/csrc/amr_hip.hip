// amr_hip.hip -- the amr_ctx context and the C ABI declared in amr_hip.h.
//
// The context owns the device state of one periodic tree (Morton keys,
// neighbour table, two compact interior buffers, regrid scratch) and one HIP
// stream.  Host code here only sequences launches and forms per-step
// constants; all per-cell and per-leaf work runs in the kernels of
// kernels_step.hpp / kernels_tree.hpp.
#include "amr_hip.h"
#include "device_common.hpp"
#include "kernels_adjoint.hpp"
#include "kernels_euler.hpp"
#include "kernels_fused.hpp"
#include "kernels_step.hpp"
#include "kernels_tree.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <fstream>
#include <limits>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

using namespace amrhip;

namespace
{

thread_local std::string g_err;

struct amr_error : std::runtime_error
{
    int code;
    amr_error(int c, std::string const& m)
        : std::runtime_error(m), code(c)
    {
    }
};

#define AMR_HIP_CHECK(expr)                                                                       \
    do                                                                                            \
    {                                                                                             \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            throw amr_error(AMR_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
    } while (0)

inline void check_launch(char const* what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw amr_error(AMR_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

constexpr int block_threads(int cells)
{
    int t = 64;
    while (t < cells && t < 256) t *= 2;
    return t;
}

// Per-launch extras of the step kernels: the +x run lists of the table
// (runs of at most kRunK[k] leaves).  A run variant without its list falls
// back to the wave-per-leaf kernel.
// rows per wave of the 2D 64^2 Euler step (k_euler_step_2d64): 16 measured
// best on the reference's 2D benchmarks (8 and 32 within 3 %, 4 12 % slower)
constexpr int      kEuler2d64Rows   = 16;
constexpr int      kRunLists        = 2;
constexpr uint32_t kRunK[kRunLists] = { 4, 8 };
struct step_aux
{
    const leaf_run* list[kRunLists]{};
    uint32_t        count[kRunLists]{};
};

// What a step-kernel variant needs besides the table.
enum : int
{
    NEEDS_RUNS = 1,
};

// Kernel table of one patch shape: (rank, extents, halo width).  Every
// function takes the table itself first, which carries the extents of the
// any-shape instance (the fixed-shape instances have them built in).
struct ops
{
    int     rank, n, h; // n: the extent of every dim of a fixed cubic shape, 0 for the any-shape instance
    int32_t ext[3];     // interior cells per layout dim
    void (*step)(ops const&, int variant, const double*, double*, const uint64_t*, const face_link*, const uint32_t*,
                 step_params const&, const step_params*, const uint16_t*, uint32_t, step_aux const&, hipStream_t);
    int (*variant_count)();
    char const* (*variant_name)(int);
    int (*variant_needs)(int);
    void (*padded)(ops const&, const double*, double*, const face_link*, const uint32_t*, uint32_t, hipStream_t);
    void (*criterion)(ops const&, int kind, const double*, const uint64_t*, const face_link*, const uint32_t*, double,
                      double, int, int, int8_t*, uint32_t, hipStream_t);
    // emit(..., guard_cap, stream): guard_cap 0 emits unconditionally; else
    // the emit checks the planned regrid on the device (regrid_applies)
    void (*emit)(ops const&, const uint64_t*, const double*, uint32_t, int, const uint8_t*, const uint8_t*,
                 const uint32_t*, const uint32_t*, uint64_t*, double*, regrid_summary*, int, size_t, uint32_t,
                 hipStream_t);
    void (*adj_step)(ops const&, const double*, const double*, const uint64_t*, const face_link*, const uint32_t*,
                     const adjoint_params*, double*, double*, uint32_t, hipStream_t);
    void (*adj_regrid)(ops const&, const double*, const uint8_t*, const uint8_t*, const uint32_t*, const uint32_t*,
                       double*, uint32_t, hipStream_t);
    // two fused steps of a single-level tree (null: not built for the shape)
    void (*step2)(const double*, double*, step2_params const&, uint32_t, hipStream_t);
    // Euler fields (field-major, stride fstride doubles between fields)
    void (*euler_dt)(ops const&, const double*, size_t, const uint64_t*, euler_params const&, euler_scalars*, uint32_t,
                     hipStream_t);
    void (*euler_step)(ops const&, const double*, double*, size_t, const uint64_t*, const face_link*, const uint32_t*,
                       euler_params const&, euler_scalars*, uint32_t, hipStream_t);

    int ni() const { return (rank == 2) ? ext[0] * ext[1] : ext[0] * ext[1] * ext[2]; }
    int padded_cells() const
    {
        int p = 1;
        for (int d = 0; d < rank; ++d) p *= ext[d] + 2 * h;
        return p;
    }
    // whether two tables run the same kernels on the same shape
    bool same_shape(ops const& o) const
    {
        return step == o.step && rank == o.rank && h == o.h && ext[0] == o.ext[0] && ext[1] == o.ext[1] &&
               ext[2] == o.ext[2];
    }
};

template <int R, int N, int H>
struct instance
{
    static constexpr int NT  = block_threads(patch_shape<R, N>::NI);
    static constexpr int P   = N + 2 * H;
    static constexpr int NTP = block_threads((R == 2) ? P * P : P * P * P);

    // Step-kernel variants (all bit-identical).  wave: one wave per leaf (the
    // default); -nt: non-temporal stores; -lds10k: 10 KiB of extra dynamic
    // LDS per workgroup, i.e. fewer leaves in flight per CU; tile256: LDS tile
    // per leaf (the generic kernel); run4 / run8: one wave walks a run of up
    // to 4 / 8 same-level leaves along +x.  The autotuner picks one per tree.
    static constexpr char const* kVariantsWave[] = { "wave", "wave-nt", "wave-lds10k", "tile256", "run4", "run8" };
    static constexpr char const* kVariantsAny[]  = { "tile" };
    static constexpr bool        kWave           = (R == 3 && N == 8) || (R == 2 && N == 16);
    static int                   variant_count() { return kWave ? 6 : 1; }
    static char const*           variant_name(int v) { return kWave ? kVariantsWave[v] : kVariantsAny[v]; }
    static int                   variant_needs(int v) { return (kWave && (v == 4 || v == 5)) ? NEEDS_RUNS : 0; }

    template <bool NTS>
    static void launch_wave(dim3 grid, unsigned lds, const double* u, double* un, const uint64_t* keys,
                            const face_link* links, const uint32_t* metas, step_params const& prm,
                            const step_params* table, const uint16_t* inst, uint32_t n, hipStream_t s)
    {
        if constexpr (R == 3 && N == 8)
            hipLaunchKernelGGL(k_step_3d8<NTS>, grid, dim3(64), lds, s, u, un, keys, links, metas, prm, table, inst, n);
        else if constexpr (R == 2 && N == 16)
            hipLaunchKernelGGL(k_step_2d16<NTS>, grid, dim3(64), lds, s, u, un, keys, links, metas, prm, table, inst, n);
    }

    static void step(ops const&, int variant, const double* u, double* un, const uint64_t* keys, const face_link* links,
                     const uint32_t* metas, step_params const& prm, const step_params* table,
                     const uint16_t* inst, uint32_t n, step_aux const& aux, hipStream_t s)
    {
        if (n == 0) return;
        const dim3 per_leaf(xcd_grid(n));
        if constexpr (kWave)
        {
            const int rk = (variant == 4 || variant == 5) ? variant - 4 : -1; // run list of the run variants
            if (rk >= 0 && !aux.list[rk]) variant = 0; // no run list: wave-per-leaf kernel
            switch (variant)
            {
                case 1: launch_wave<true>(per_leaf, 0, u, un, keys, links, metas, prm, table, inst, n, s); break;
                case 2: launch_wave<false>(per_leaf, 10 << 10, u, un, keys, links, metas, prm, table, inst, n, s); break;
                case 3:
                    hipLaunchKernelGGL((k_step<R, N, NT>), per_leaf, dim3(NT), 0, s, u, un, keys, links, metas, prm,
                                       table, inst, n);
                    break;
                case 4:
                case 5:
                {
                    const dim3 per_run(xcd_grid(aux.count[rk]));
                    if constexpr (R == 3)
                        hipLaunchKernelGGL(k_step_3d8_run<false>, per_run, dim3(64), 0, s, u, un, keys, links, metas,
                                           prm, table, inst, aux.list[rk], aux.count[rk]);
                    else
                        hipLaunchKernelGGL(k_step_2d16_run<false>, per_run, dim3(64), 0, s, u, un, keys, links,
                                           metas, prm, table, inst, aux.list[rk], aux.count[rk]);
                    break;
                }
                default: launch_wave<false>(per_leaf, 0, u, un, keys, links, metas, prm, table, inst, n, s);
            }
        }
        else
            hipLaunchKernelGGL((k_step<R, N, NT>), per_leaf, dim3(NT), 0, s, u, un, keys, links, metas, prm, table,
                               inst, n);
        check_launch("k_step");
    }
    static void padded(ops const&, const double* u, double* pad, const face_link* links, const uint32_t* metas, uint32_t n,
                       hipStream_t s)
    {
        if (n == 0) return;
        hipLaunchKernelGGL((k_padded_view<R, N, H, NTP>), dim3(n), dim3(NTP), 0, s, u, pad, links, metas, n);
        check_launch("k_padded_view");
    }
    static void criterion(ops const&, int kind, const double* u, const uint64_t* keys, const face_link* links,
                          const uint32_t* metas, double r, double c, int mn, int mx, int8_t* flags, uint32_t n,
                          hipStream_t s)
    {
        if (n == 0) return;
        if (kind == AMR_CRITERION_GRADIENT)
            hipLaunchKernelGGL((k_criterion_gradient<R, N, NT>), dim3(n), dim3(NT), 0, s, u, keys, links, metas, r, c,
                               mn, mx, flags, n);
        else
            hipLaunchKernelGGL((k_criterion_max<R, N, H, NTP>), dim3(n), dim3(NTP), 0, s, u, keys, links, metas, r, c,
                               mn, mx, flags, n);
        check_launch("k_criterion");
    }
    static void emit(ops const&, const uint64_t* keys, const double* u, uint32_t n, int depth, const uint8_t* mark,
                     const uint8_t* head, const uint32_t* counts, const uint32_t* pos, uint64_t* keys_out,
                     double* u_out, regrid_summary* summary, int nfields, size_t fstride, uint32_t guard_cap,
                     hipStream_t s)
    {
        if (n == 0) return;
        hipLaunchKernelGGL((k_emit<R, N, NT>), dim3(xcd_grid(n)), dim3(NT), 0, s, keys, u, n, depth, mark, head, counts,
                           pos, keys_out, u_out, summary, nfields, fstride, guard_cap);
        check_launch("k_emit");
    }
    static void adj_step(ops const&, const double* lam_next, const double* u, const uint64_t* keys, const face_link* links,
                         const uint32_t* metas, const adjoint_params* prm, double* lam, double* vpart, uint32_t n,
                         hipStream_t s)
    {
        if (n == 0) return;
        hipLaunchKernelGGL((k_adj_step<R, N, NT>), dim3(n), dim3(NT), 0, s, lam_next, u, keys, links, metas, prm, lam,
                           vpart, n);
        check_launch("k_adj_step");
    }
    static void adj_regrid(ops const&, const double* lam_new, const uint8_t* mark, const uint8_t* head, const uint32_t* counts,
                           const uint32_t* pos, double* lam_old, uint32_t n_old, hipStream_t s)
    {
        if (n_old == 0) return;
        hipLaunchKernelGGL((k_adj_regrid<R, N, NT>), dim3(n_old), dim3(NT), 0, s, lam_new, mark, head, counts, pos,
                           lam_old, n_old);
        check_launch("k_adj_regrid");
    }
    static void step2(const double* u, double* un, step2_params const& prm, uint32_t n, hipStream_t s)
    {
        if (n == 0) return;
        if constexpr (R == 3 && N == 8)
        {
            // levels >= 1: a workgroup per 2 x 2 x 1 block of sibling leaves
            // (n = 8^level, so the blocks tile the slots); the root alone:
            // one wave
            if (n % 4 == 0)
                hipLaunchKernelGGL((k_step2_3d8_uniform<false, 4>), dim3(xcd_grid(n / 4)), dim3(256), 0, s, u, un, prm,
                                   n);
            else
                hipLaunchKernelGGL((k_step2_3d8_uniform<false, 1>), dim3(xcd_grid(n)), dim3(64), 0, s, u, un, prm, n);
        }
        check_launch("k_step2");
    }
    static void euler_dt(ops const&, const double* u, size_t fs, const uint64_t* keys, euler_params const& prm, euler_scalars* sc,
                         uint32_t n, hipStream_t s)
    {
        if (n == 0) return;
        hipLaunchKernelGGL((k_euler_dt<R, N, NT>), dim3(xcd_grid(n)), dim3(NT), 0, s, u, fs, keys, prm, sc, n);
        check_launch("k_euler_dt");
    }
    static void euler_step(ops const&, const double* u, double* un, size_t fs, const uint64_t* keys, const face_link* links,
                           const uint32_t* metas, euler_params const& prm, euler_scalars* sc, uint32_t n,
                           hipStream_t s)
    {
        if (n == 0) return;
        // 3D 8^3: one wave per leaf with the column in registers; other
        // shapes: LDS tiles of every field where they fit, else LDS strips of
        // the patch, else neighbours from memory
        constexpr size_t tile_bytes = (R + 2) * tile_shape<R, N>::TS * sizeof(double);
        constexpr int    B          = euler_strip_depth<R, N>(40 * 1024);
        if constexpr (R == 3 && N == 8)
            hipLaunchKernelGGL((k_euler_step_3d8<false>), dim3(xcd_grid(n)), dim3(64), 0, s, u, un, fs, keys, links,
                               metas, prm, sc, n);
        else if constexpr (R == 2 && N == 64)
            hipLaunchKernelGGL((k_euler_step_2d64<kEuler2d64Rows>), dim3(xcd_grid(static_cast<uint64_t>(n) * (64 / kEuler2d64Rows))),
                               dim3(64), 0, s, u, un, fs, keys, links, metas, prm, sc, n);
        else if constexpr (tile_bytes <= 48 * 1024)
            hipLaunchKernelGGL((k_euler_step_tile<R, N, NT>), dim3(xcd_grid(n)), dim3(NT), 0, s, u, un, fs, keys, links,
                               metas, prm, sc, n);
        else if constexpr (B > 0)
            hipLaunchKernelGGL((k_euler_step_strip<R, N, B, NT>), dim3(xcd_grid(static_cast<uint64_t>(n) * (N / B))),
                               dim3(NT), 0, s, u, un, fs, keys, links, metas, prm, sc, n);
        else
            hipLaunchKernelGGL((k_euler_step<R, N, NT>), dim3(xcd_grid(n)), dim3(NT), 0, s, u, un, fs, keys, links,
                               metas, prm, sc, n);
        check_launch("k_euler_step");
    }
    static constexpr ops table()
    {
        constexpr bool fused = R == 3 && N == 8;
        return ops{ R,
                    N,
                    H,
                    { N, N, R == 3 ? N : 1 },
                    &step,
                    &variant_count,
                    &variant_name,
                    &variant_needs,
                    &padded,
                    &criterion,
                    &emit,
                    &adj_step,
                    &adj_regrid,
                    fused ? &step2 : nullptr,
                    &euler_dt,
                    &euler_step };
    }
};

// Patch shapes built into this library: the configurations' 16^2 and 8^3
// patches with one ghost layer, the reference test's 10^2 and 50^2 patches
// with two, the reference's 2D Euler benchmark's 64^2 patches
// (benchmark/bench_fvm_solver_integration.b.cpp:35-37), and small shapes for
// exhaustive parity tests.
const ops kOps[] = {
    instance<2, 16, 1>::table(), instance<3, 8, 1>::table(), instance<2, 10, 2>::table(),
    instance<2, 50, 2>::table(), instance<2, 4, 1>::table(), instance<2, 4, 2>::table(),
    instance<2, 8, 1>::table(),  instance<3, 4, 1>::table(), instance<2, 64, 1>::table(),
};

ops const* find_ops(int rank, int n, int h)
{
    for (auto const& o : kOps)
        if (o.rank == rank && o.n == n && o.h == h) return &o;
    return nullptr;
}

// Kernels for every other patch shape (any even extents, square or not):
// the box kernels, the extents passed at launch, one workgroup per leaf of
// 64 threads for patches of up to 64 cells and 256 otherwise.  They read
// neighbour cells from memory instead of LDS tiles; the arithmetic is the
// fixed-shape kernels', so results are the same bits.
template <int R>
struct box_instance
{
    static box<R> shape(ops const& o) { return box<R>{ { o.ext[0], o.ext[1], o.ext[2] } }; }
    static bool   small(ops const& o) { return o.ni() <= 64; }

    static int         variant_count() { return 1; }
    static char const* variant_name(int) { return "box"; }
    static int         variant_needs(int) { return 0; }

    static void step(ops const& o, int, const double* u, double* un, const uint64_t* keys, const face_link* links,
                     const uint32_t* metas, step_params const& prm, const step_params* table, const uint16_t* inst,
                     uint32_t n, step_aux const&, hipStream_t s)
    {
        if (n == 0) return;
        const dim3 grid(xcd_grid(n));
        if (small(o))
            hipLaunchKernelGGL((k_step_box<R, 64>), grid, dim3(64), 0, s, shape(o), u, un, keys, links, metas, prm,
                               table, inst, n);
        else
            hipLaunchKernelGGL((k_step_box<R, 256>), grid, dim3(256), 0, s, shape(o), u, un, keys, links, metas, prm,
                               table, inst, n);
        check_launch("k_step_box");
    }
    static void padded(ops const& o, const double* u, double* pad, const face_link* links, const uint32_t* metas,
                       uint32_t n, hipStream_t s)
    {
        if (n == 0) return;
        hipLaunchKernelGGL((k_padded_box<R, 256>), dim3(n), dim3(256), 0, s, shape(o), o.h, u, pad, links, metas, n);
        check_launch("k_padded_box");
    }
    static void criterion(ops const& o, int kind, const double* u, const uint64_t* keys, const face_link* links,
                          const uint32_t* metas, double r, double c, int mn, int mx, int8_t* flags, uint32_t n,
                          hipStream_t s)
    {
        if (n == 0) return;
        if (kind == AMR_CRITERION_GRADIENT)
            hipLaunchKernelGGL((k_criterion_gradient_box<R, 256>), dim3(n), dim3(256), 0, s, shape(o), u, keys, links,
                               metas, r, c, mn, mx, flags, n);
        else
            hipLaunchKernelGGL((k_criterion_max_box<R, 256>), dim3(n), dim3(256), 0, s, shape(o), o.h, u, keys, links,
                               metas, r, c, mn, mx, flags, n);
        check_launch("k_criterion_box");
    }
    static void emit(ops const& o, const uint64_t* keys, const double* u, uint32_t n, int depth, const uint8_t* mark,
                     const uint8_t* head, const uint32_t* counts, const uint32_t* pos, uint64_t* keys_out,
                     double* u_out, regrid_summary* summary, int nfields, size_t fstride, uint32_t guard_cap,
                     hipStream_t s)
    {
        if (n == 0) return;
        hipLaunchKernelGGL((k_emit_box<R, 256>), dim3(xcd_grid(n)), dim3(256), 0, s, shape(o), keys, u, n, depth, mark,
                           head, counts, pos, keys_out, u_out, summary, nfields, fstride, guard_cap);
        check_launch("k_emit_box");
    }
    static void adj_step(ops const& o, const double* lam_next, const double* u, const uint64_t* keys,
                         const face_link* links, const uint32_t* metas, const adjoint_params* prm, double* lam,
                         double* vpart, uint32_t n, hipStream_t s)
    {
        if (n == 0) return;
        hipLaunchKernelGGL((k_adj_step_box<R, 256>), dim3(n), dim3(256), 0, s, shape(o), lam_next, u, keys, links,
                           metas, prm, lam, vpart, n);
        check_launch("k_adj_step_box");
    }
    static void adj_regrid(ops const& o, const double* lam_new, const uint8_t* mark, const uint8_t* head,
                           const uint32_t* counts, const uint32_t* pos, double* lam_old, uint32_t n_old, hipStream_t s)
    {
        if (n_old == 0) return;
        hipLaunchKernelGGL((k_adj_regrid_box<R, 256>), dim3(n_old), dim3(256), 0, s, shape(o), lam_new, mark, head,
                           counts, pos, lam_old, n_old);
        check_launch("k_adj_regrid_box");
    }
    static void euler_dt(ops const& o, const double* u, size_t fs, const uint64_t* keys, euler_params const& prm,
                         euler_scalars* sc, uint32_t n, hipStream_t s)
    {
        if (n == 0) return;
        hipLaunchKernelGGL((k_euler_dt_box<R, 256>), dim3(xcd_grid(n)), dim3(256), 0, s, shape(o), u, fs, keys, prm, sc,
                           n);
        check_launch("k_euler_dt_box");
    }
    static void euler_step(ops const& o, const double* u, double* un, size_t fs, const uint64_t* keys,
                           const face_link* links, const uint32_t* metas, euler_params const& prm, euler_scalars* sc,
                           uint32_t n, hipStream_t s)
    {
        if (n == 0) return;
        hipLaunchKernelGGL((k_euler_step_box<R, 256>), dim3(xcd_grid(n)), dim3(256), 0, s, shape(o), u, un, fs, keys,
                           links, metas, prm, sc, n);
        check_launch("k_euler_step_box");
    }
    static ops table(int h, const int32_t* ext)
    {
        return ops{ R,       0,          h,     { ext[0], ext[1], R == 3 ? ext[2] : 1 },
                    &step,   &variant_count, &variant_name, &variant_needs,
                    &padded, &criterion, &emit, &adj_step,
                    &adj_regrid, nullptr, &euler_dt, &euler_step };
    }
};

// The kernel table of a patch shape: a fixed-shape instance for the cubic
// shapes built above, else the any-shape kernels.
ops shape_ops(int rank, const int32_t* ext, int h)
{
    bool cubic = true;
    for (int d = 1; d < rank; ++d) cubic = cubic && ext[d] == ext[0];
    if (ops const* o = cubic ? find_ops(rank, ext[0], h) : nullptr) return *o;
    return rank == 2 ? box_instance<2>::table(h, ext) : box_instance<3>::table(h, ext);
}

template <typename Fn>
int guarded(Fn&& fn)
{
    try
    {
        fn();
        g_err.clear();
        return AMR_OK;
    }
    catch (amr_error const& e)
    {
        g_err = e.what();
        return e.code;
    }
    catch (std::bad_alloc const&)
    {
        g_err = "host allocation failed";
        return AMR_ERR_CAPACITY;
    }
    catch (std::exception const& e)
    {
        g_err = e.what();
        return AMR_ERR_ARGUMENT;
    }
}

template <typename T>
T* dev_alloc(size_t count)
{
    void* p = nullptr;
    AMR_HIP_CHECK(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
    return static_cast<T*>(p);
}

inline unsigned grid1d(uint64_t items, unsigned threads) { return static_cast<unsigned>((items + threads - 1) / threads); }

// Times `trials` launches of every variant, three interleaved rounds, and
// returns the variant with the lowest per-launch time (min over rounds).
template <typename Launch>
int pick_fastest(int nvar, Launch&& launch, hipStream_t stream, uint32_t trials, float* best_us)
{
    hipEvent_t e0, e1;
    AMR_HIP_CHECK(hipEventCreate(&e0));
    AMR_HIP_CHECK(hipEventCreate(&e1));
    std::vector<float> best(static_cast<size_t>(nvar), std::numeric_limits<float>::max());
    trials = std::max(trials, 1u);
    for (int round = 0; round < 3; ++round)
    {
        for (int v = 0; v < nvar; ++v)
        {
            launch(v); // warm
            AMR_HIP_CHECK(hipEventRecord(e0, stream));
            for (uint32_t t = 0; t < trials; ++t) launch(v);
            AMR_HIP_CHECK(hipEventRecord(e1, stream));
            AMR_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0.0f;
            AMR_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            best[static_cast<size_t>(v)] = std::min(best[static_cast<size_t>(v)], 1000.0f * ms / static_cast<float>(trials));
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const auto it = std::min_element(best.begin(), best.end());
    if (best_us) *best_us = *it;
    return static_cast<int>(it - best.begin());
}

// Exclusive scan of n uint32 values into out; the total lands in
// summary->total.
void exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* block_sums, regrid_summary* summary,
                    hipStream_t s)
{
    const unsigned nb = std::max(grid1d(n, kScanBlock), 1u);
    hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(kScanThreads), 0, s, in, out, n, block_sums);
    check_launch("k_scan_blocks");
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanThreads), 0, s, block_sums, nb, summary);
    check_launch("k_scan_sums");
    if (n == 0) return;
    hipLaunchKernelGGL(k_scan_add, dim3(grid1d(n, 256)), dim3(256), 0, s, out, n, block_sums);
    check_launch("k_scan_add");
}

// The +x run lists (runs of at most 4 and 8 leaves) of one leaf table, for
// the shapes that have run kernels.
struct run_lists
{
    leaf_run*       list[kRunLists]{};
    uint32_t        count[kRunLists]{};
    uint32_t*       flag       = nullptr;
    uint32_t*       pos        = nullptr;
    uint32_t*       block_sums = nullptr;
    regrid_summary* summary    = nullptr;
    regrid_summary* h_summary  = nullptr; // pinned, one per list
    bool            valid      = false;   // built for the current table

    run_lists() = default;
    run_lists(run_lists const&) = delete;
    run_lists& operator=(run_lists const&) = delete;
    ~run_lists()
    {
        for (int k = 0; k < kRunLists; ++k)
            if (list[k]) (void)hipFree(list[k]);
        for (void* p : { (void*)flag, (void*)pos, (void*)block_sums, (void*)summary })
            if (p) (void)hipFree(p);
        if (h_summary) (void)hipHostFree(h_summary);
    }

    static bool wanted(ops const* op) { return (op->rank == 3 && op->n == 8) || (op->rank == 2 && op->n == 16); }

    void allocate(uint64_t cap)
    {
        for (int k = 0; k < kRunLists; ++k) list[k] = dev_alloc<leaf_run>(cap);
        flag       = dev_alloc<uint32_t>(cap);
        pos        = dev_alloc<uint32_t>(cap);
        block_sums = dev_alloc<uint32_t>(cap / kScanBlock + 2);
        summary    = dev_alloc<regrid_summary>(kRunLists);
        AMR_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&h_summary), kRunLists * sizeof(regrid_summary), 0));
    }

    // The lists for a step launch; none while they are stale (a run variant
    // then runs the wave-per-leaf kernel).
    step_aux view() const
    {
        step_aux v;
        if (!valid) return v;
        for (int k = 0; k < kRunLists; ++k)
        {
            v.list[k]  = list[k];
            v.count[k] = count[k];
        }
        return v;
    }

    // Builds both lists for a valid table of n leaves and waits for them.
    void build(int rank, const uint64_t* keys, uint32_t n, int depth, const face_link* links, const uint32_t* metas,
               hipStream_t s)
    {
        if (!flag) return;
        AMR_HIP_CHECK(hipMemsetAsync(summary, 0, kRunLists * sizeof(regrid_summary), s));
        for (int k = 0; k < kRunLists; ++k)
        {
            if (rank == 2)
                hipLaunchKernelGGL(k_run_heads<2>, dim3(grid1d(n, 256)), dim3(256), 0, s, keys, n, depth, metas, kRunK[k], flag);
            else
                hipLaunchKernelGGL(k_run_heads<3>, dim3(grid1d(n, 256)), dim3(256), 0, s, keys, n, depth, metas, kRunK[k], flag);
            check_launch("k_run_heads");
            exclusive_scan(flag, pos, n, block_sums, summary + k, s);
            if (rank == 2)
                hipLaunchKernelGGL(k_run_emit<2>, dim3(grid1d(n, 256)), dim3(256), 0, s, keys, n, depth, links, metas,
                                   kRunK[k], flag, pos, list[k]);
            else
                hipLaunchKernelGGL(k_run_emit<3>, dim3(grid1d(n, 256)), dim3(256), 0, s, keys, n, depth, links, metas,
                                   kRunK[k], flag, pos, list[k]);
            check_launch("k_run_emit");
        }
        AMR_HIP_CHECK(hipMemcpyAsync(h_summary, summary, kRunLists * sizeof(regrid_summary), hipMemcpyDeviceToHost, s));
        AMR_HIP_CHECK(hipStreamSynchronize(s));
        for (int k = 0; k < kRunLists; ++k) count[k] = h_summary[k].total;
        valid = true;
    }
};

// Times every step-kernel variant on trial steps that evolve a copy of the
// state u_cur, ping-ponging between u_other and a temporary buffer, so that
// every launch reads what the previous one wrote (as in a solve); u_cur is
// not modified.  launch(v, src, dst, aux) queues one step.  Returns the
// fastest variant.
template <typename Launch>
int autotune_steps(ops const* op, const double* u_cur, double* u_other, uint64_t leaves, int ni, Launch&& launch,
                   hipStream_t stream, uint32_t trials, float* best_us)
{
    double* tmp  = dev_alloc<double>(leaves * ni);
    int     best = 0;
    try
    {
        double* bufs[2] = { tmp, u_other };
        int     p       = 0;
        AMR_HIP_CHECK(hipMemcpyAsync(tmp, u_cur, leaves * ni * sizeof(double), hipMemcpyDeviceToDevice, stream));
        best = pick_fastest(
            op->variant_count(),
            [&](int v)
            {
                launch(v, bufs[p], bufs[1 - p]);
                p = 1 - p;
            },
            stream, trials, best_us);
    }
    catch (...)
    {
        (void)hipStreamSynchronize(stream);
        (void)hipFree(tmp);
        throw;
    }
    AMR_HIP_CHECK(hipStreamSynchronize(stream));
    AMR_HIP_CHECK(hipFree(tmp));
    return best;
}

// Batch packing: shift an instance's neighbour slots by its leaf offset, and
// tag its leaves with the instance index.
__global__ void k_offset_links(face_link* __restrict__ links, uint64_t count, int32_t offset)
{
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i >= count) return;
    for (int j = 0; j < 4; ++j)
        if (links[i].ids[j] >= 0) links[i].ids[j] += offset;
}

__global__ void k_fill_instance(uint16_t* __restrict__ inst, uint64_t count, uint16_t value)
{
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i < count) inst[i] = value;
}

} // namespace

// A context's read-back block in pinned host memory.  k_publish, queued
// behind the work whose results the host needs, copies the regrid summary
// and / or the Euler batch scalars into it and then stores the sequence
// number with a system-scope release, so the host learns the outcome by
// reading host memory instead of synchronizing the stream.
struct host_mailbox
{
    regrid_summary summary;
    euler_scalars  esc;
    uint32_t       seq;
    uint32_t       pad[3];
};

__global__ void k_publish(const regrid_summary* __restrict__ s, const euler_scalars* __restrict__ e,
                          host_mailbox* box, uint32_t seq)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (s) box->summary = *s;
    if (e) box->esc = *e;
    __threadfence_system();
    __hip_atomic_store(&box->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// How long the host spins on the mailbox before it falls back to a stream
// synchronization: waits for a regrid or a short batch of steps end within
// it (a spin sees the outcome about 8 us sooner than a synchronization).
constexpr int64_t kMailboxSpinUs = 200;

struct amr_ctx
{
    amr_tree_config cfg{};
    ops             op_table{};     // this tree's kernel table (its patch shape)
    ops const*      op = &op_table;
    hipStream_t     stream{};
    int             device = 0; // the GPU current at creation (owns the stream and buffers)
    int             rank = 0, depth = 0, ni = 0, flat = 0, ndir = 0;
    uint64_t        cap = 0;

    // topology + state
    uint64_t        n         = 0;
    int             max_level = 0;
    int             cur       = 0; // current interior buffer (flips every step and regrid)
    int             kc        = 0; // current key buffer (flips every regrid)
    double*         u[2]{};
    uint64_t*       keys[2]{};
    face_link*      links = nullptr;
    uint32_t*       metas = nullptr;
    double*         padded = nullptr; // lazily allocated padded view

    // regrid scratch
    int8_t*         flags = nullptr;
    uint8_t*        mark = nullptr;
    uint8_t*        head = nullptr;
    uint32_t*       counts = nullptr;
    uint32_t*       pos = nullptr;
    uint32_t*       block_sums = nullptr;
    regrid_summary* summary = nullptr;
    host_mailbox*   box     = nullptr; // pinned read-back block
    uint32_t        box_seq = 0;

    // solver
    double          vel[3] = { 1.0, 0.5, 0.0 };
    double          cfl    = 0.4;
    double          pending_dt    = 0.0;
    uint64_t        pending_steps = 0;
    bool            broken        = false;
    int             variant       = 0; // step-kernel variant (autotuner)
    run_lists       runs;              // +x runs of the current table
    uint64_t        sweeps        = 0; // step launches (sets the sweep direction)
    bool            fuse          = true; // two steps per launch on single-level trees

    // fields: 1 (advection) or rank + 2 conservative Euler variables,
    // field-major (field v of slot s at u + v * fstride() + s * ni)
    int             nvar  = 1;
    double          gamma = 1.4;
    euler_scalars*  esc   = nullptr; // device scalars of an Euler batch
    size_t          fstride() const { return static_cast<size_t>(cap) * static_cast<size_t>(ni); }
    bool            euler() const { return nvar > 1; }

    ~amr_ctx()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void* p : { (void*)u[0], (void*)u[1], (void*)keys[0], (void*)keys[1], (void*)links, (void*)metas,
                         (void*)padded, (void*)flags, (void*)mark, (void*)head, (void*)counts, (void*)pos,
                         (void*)block_sums, (void*)summary, (void*)esc })
            if (p) (void)hipFree(p);
        if (box) (void)hipHostFree(box);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // ---- geometry (include/solver/physics_system.hpp:150-178) ----------
    double cell_size(int level, int axis) const
    {
        const double pm   = static_cast<double>(1u << (depth - level));
        const double maxc = static_cast<double>(1u << depth);
        const double ps   = cfg.domain_length[axis] * pm / maxc;
        return ps / static_cast<double>(op->ext[rank - 1 - axis]); // cells along physical axis `axis`
    }

    // compute_time_step_cpu (amr_solver.hpp:355-413): the per-cell minimum is
    // attained on the finest level present.
    double compute_dt() const
    {
        double dt = std::numeric_limits<double>::max();
        for (int d = 0; d < rank; ++d)
        {
            const double speed = std::abs(vel[d]);
            if (speed > 1e-12) dt = std::min(dt, cell_size(max_level, d) / speed);
        }
        return cfl * dt;
    }

    void sync() { AMR_HIP_CHECK(hipStreamSynchronize(stream)); }

    // Waits until every operation queued on the stream so far has finished
    // and the mailbox holds the given summary / Euler scalars (either may be
    // null): spins on the mailbox's sequence number for up to
    // kMailboxSpinUs, then synchronizes the stream (which also reports a
    // failure of the queued work).
    void publish_wait(const regrid_summary* s, const euler_scalars* e)
    {
        const uint32_t want = ++box_seq;
        hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, stream, s, e, box, want);
        check_launch("k_publish");
        const auto t0 = std::chrono::steady_clock::now();
        for (uint32_t i = 1;; ++i)
        {
            if (__atomic_load_n(&box->seq, __ATOMIC_ACQUIRE) == want) return;
            if ((i & 255u) == 0u &&
                std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(kMailboxSpinUs))
                break;
        }
        sync();
        if (__atomic_load_n(&box->seq, __ATOMIC_ACQUIRE) != want)
            throw amr_error(AMR_ERR_HIP, "the read-back of a finished stream did not arrive");
    }

    // Neighbour records of the current leaves (a regrid builds its new table
    // itself, see regrid).
    void build_table()
    {
        AMR_HIP_CHECK(hipMemsetAsync(summary, 0, sizeof(regrid_summary), stream));
        const uint64_t items = n * static_cast<uint64_t>(ndir);
        if (rank == 2)
            hipLaunchKernelGGL(k_build_table<2>, dim3(grid1d(items, 256)), dim3(256), 0, stream, keys[kc],
                               static_cast<uint32_t>(n), depth, links, metas, summary, 0u);
        else
            hipLaunchKernelGGL(k_build_table<3>, dim3(grid1d(items, 256)), dim3(256), 0, stream, keys[kc],
                               static_cast<uint32_t>(n), depth, links, metas, summary, 0u);
        check_launch("k_build_table");
        publish_wait(summary, nullptr);
        if (box->summary.err & 1u)
        {
            // the table may hold -1 entries: no kernel may run on it again
            broken = true;
            throw amr_error(AMR_ERR_TOPOLOGY, "leaves do not tile the domain or violate 2:1 face balance");
        }
        // the run lists are rebuilt when a step or the autotuner next needs
        // them (most trees run the wave-per-leaf kernel, which does not)
        runs.valid = false;
    }

    // Builds the run lists of the current table if `v` (or, v < 0, any
    // variant) needs them and they are stale.  Synchronizes the stream: not
    // for use inside a graph capture.
    void ensure_runs(int v)
    {
        if (runs.valid || !runs.flag) return;
        if (v >= 0 && !(op->variant_needs(v) & NEEDS_RUNS)) return;
        runs.build(rank, keys[kc], static_cast<uint32_t>(n), depth, links, metas, stream);
    }

    void ensure_usable() const
    {
        if (broken) throw amr_error(AMR_ERR_TOPOLOGY, "context unusable after a topology error");
    }

    step_params make_params(double dt) const
    {
        step_params prm{};
        for (int d = 0; d < 3; ++d)
        {
            prm.vel[d] = vel[d];
            prm.hs[d]  = 0.5 * std::abs(vel[d]);
        }
        for (int l = 0; l <= depth; ++l)
            for (int d = 0; d < rank; ++d) prm.dtdx[l][d] = dt / cell_size(l, d);
        return prm;
    }

    // Whether the next two steps can run as one fused launch: the shape has
    // the fused kernel and every leaf sits on the finest level (the leaves
    // tile the domain, so n == 2^(rank * max_level) holds exactly then).
    bool fusion_applies() const
    {
        return nvar == 1 && op->step2 && rank * max_level < 32 && n == (uint64_t{ 1 } << (rank * max_level));
    }
    bool fusable() const { return fuse && fusion_applies(); }

    // Two steps (dt1 then dt2) in one launch; bit-identical to step(dt1),
    // step(dt2).
    void step2(double dt1, double dt2)
    {
        ensure_usable();
        step2_params prm{};
        for (int d = 0; d < 3; ++d)
        {
            prm.vel[d] = vel[d];
            prm.hs[d]  = 0.5 * std::abs(vel[d]);
        }
        for (int d = 0; d < rank; ++d)
        {
            prm.dtdx[0][d] = dt1 / cell_size(max_level, d);
            prm.dtdx[1][d] = dt2 / cell_size(max_level, d);
        }
        prm.level   = static_cast<uint32_t>(max_level);
        prm.reverse = static_cast<uint32_t>(sweeps++ & 1);
        op->step2(u[cur], u[1 - cur], prm, static_cast<uint32_t>(n), stream);
        cur = 1 - cur;
    }

    void step(double dt)
    {
        ensure_usable();
        ensure_runs(variant);
        step_params prm = make_params(dt);
        prm.reverse     = static_cast<uint32_t>(sweeps++ & 1);
        op->step(*op, variant, u[cur], u[1 - cur], keys[kc], links, metas, prm, nullptr, nullptr, static_cast<uint32_t>(n),
                 runs.view(), stream);
        cur = 1 - cur;
    }

    // Times every step-kernel variant on the current tree (trial steps on a
    // copy of the state, see autotune_steps) and keeps the fastest.
    int autotune(uint32_t trials, float* best_us)
    {
        ensure_usable();
        ensure_runs(-1);
        step_params prm = make_params(compute_dt());
        variant         = autotune_steps(
            op, u[cur], u[1 - cur], n, ni,
            [&](int v, const double* src, double* dst)
            {
                prm.reverse ^= 1u; // trial steps alternate like a solve's
                op->step(*op, v, src, dst, keys[kc], links, metas, prm, nullptr, nullptr, static_cast<uint32_t>(n),
                         runs.view(), stream);
            },
            stream, trials, best_us);
        return variant;
    }

    // ---- Euler (EulerPhysics<R>, amr_solver.hpp:155-413) -----------------
    euler_params make_euler_params() const
    {
        euler_params prm{};
        prm.gamma = gamma;
        for (int l = 0; l <= depth; ++l)
            for (int d = 0; d < rank; ++d) prm.cs[l][d] = cell_size(l, d);
        return prm;
    }

    // advance_batch_async: the dt reduction of the batch's first state, then
    // per step the dt logic and the update, which also folds the next
    // step's dt minimum over the cells it writes (the values the reduction
    // would read back, so the same minimum); dt never visits the host.
    void euler_advance(uint64_t steps, double remaining)
    {
        ensure_usable();
        euler_params prm = make_euler_params();
        hipLaunchKernelGGL(k_euler_begin, dim3(1), dim3(1), 0, stream, esc, remaining);
        check_launch("k_euler_begin");
        const uint32_t nn = static_cast<uint32_t>(n);
        if (steps > 0) op->euler_dt(*op, u[cur], fstride(), keys[kc], prm, esc, nn, stream);
        for (uint64_t s = 0; s < steps; ++s)
        {
            hipLaunchKernelGGL(k_euler_finalize, dim3(1), dim3(1), 0, stream, esc, cfl);
            check_launch("k_euler_finalize");
            prm.reverse = static_cast<uint32_t>(sweeps++ & 1);
            prm.next_dt = s + 1 < steps ? 1u : 0u;
            op->euler_step(*op, u[cur], u[1 - cur], fstride(), keys[kc], links, metas, prm, esc, nn, stream);
            cur = 1 - cur;
        }
    }

    // The Euler batch scalars once every queued operation has finished.
    euler_scalars euler_read()
    {
        publish_wait(nullptr, esc);
        return box->esc;
    }

    // compute_time_step_cpu: cfl * min over cells and dims of dx / speed.
    double euler_dt_now()
    {
        ensure_usable();
        const euler_params prm = make_euler_params();
        hipLaunchKernelGGL(k_euler_begin, dim3(1), dim3(1), 0, stream, esc, 0.0);
        check_launch("k_euler_begin");
        op->euler_dt(*op, u[cur], fstride(), keys[kc], prm, esc, static_cast<uint32_t>(n), stream);
        const euler_scalars e = euler_read();
        double raw;
        std::memcpy(&raw, &e.dtmin_bits, sizeof(raw));
        return cfl * raw;
    }

    void scan(uint32_t count) { exclusive_scan(counts, pos, count, block_sums, summary, stream); }

    // The regrid's planning passes as one launch per pass (any tree size):
    // refine marks, the closure per level, coarsening families, output
    // counts, their scan.
    void plan_regrid(uint32_t nn)
    {
        hipLaunchKernelGGL(k_mark_refine, dim3(grid1d(nn, 256)), dim3(256), 0, stream, flags, keys[kc], nn, depth, mark);
        check_launch("k_mark_refine");
        for (int l = std::min(max_level, depth); l >= 1; --l)
        {
            if (rank == 2)
                hipLaunchKernelGGL(k_close_level<2>, dim3(grid1d(nn, 256)), dim3(256), 0, stream, keys[kc], nn, l,
                                   links, metas, mark);
            else
                hipLaunchKernelGGL(k_close_level<3>, dim3(grid1d(nn, 256)), dim3(256), 0, stream, keys[kc], nn, l,
                                   links, metas, mark);
            check_launch("k_close_level");
        }
        if (rank == 2)
        {
            hipLaunchKernelGGL(k_coarsen_families<2>, dim3(grid1d(nn, 256)), dim3(256), 0, stream, flags, keys[kc], nn,
                               depth, links, metas, mark, head, summary);
            hipLaunchKernelGGL(k_emit_counts<2>, dim3(grid1d(nn, 256)), dim3(256), 0, stream, keys[kc], nn, depth,
                               mark, head, counts, summary);
        }
        else
        {
            hipLaunchKernelGGL(k_coarsen_families<3>, dim3(grid1d(nn, 256)), dim3(256), 0, stream, flags, keys[kc], nn,
                               depth, links, metas, mark, head, summary);
            hipLaunchKernelGGL(k_emit_counts<3>, dim3(grid1d(nn, 256)), dim3(256), 0, stream, keys[kc], nn, depth,
                               mark, head, counts, summary);
        }
        check_launch("k_coarsen_families / k_emit_counts");
        scan(nn);
    }

    // ndtree::reconstruct_tree (ndtree.hpp:1249-1271) with the flags already
    // in device memory.  The plan, the emit and the new table are queued
    // back to back: the emit and the table build read the plan's summary on
    // the device and do nothing unless the regrid changes the tree, asks for
    // no root coarsening and fits the capacity (regrid_applies), so the host
    // reads the outcome once, after all of them.  The links and metas are
    // overwritten only when the regrid applies; the state and keys are
    // written to the other buffers, which become current then.
    bool regrid()
    {
        ensure_usable();
        const uint32_t nn  = static_cast<uint32_t>(n);
        const uint32_t cap32 = static_cast<uint32_t>(std::min<uint64_t>(cap, 0xffffffffu));
        if (nn <= static_cast<uint32_t>(kPlanMaxLeaves))
        {
            // small tree: every planning pass in one workgroup, which also
            // clears the summary
            const int top = std::min(max_level, depth);
            if (rank == 2)
                hipLaunchKernelGGL(k_regrid_plan<2>, dim3(1), dim3(kPlanThreads), 0, stream, flags, keys[kc], nn,
                                   depth, top, links, metas, mark, head, counts, pos, summary);
            else
                hipLaunchKernelGGL(k_regrid_plan<3>, dim3(1), dim3(kPlanThreads), 0, stream, flags, keys[kc], nn,
                                   depth, top, links, metas, mark, head, counts, pos, summary);
            check_launch("k_regrid_plan");
        }
        else
        {
            AMR_HIP_CHECK(hipMemsetAsync(summary, 0, sizeof(regrid_summary), stream));
            plan_regrid(nn);
        }
        op->emit(*op, keys[kc], u[cur], nn, depth, mark, head, counts, pos, keys[1 - kc], u[1 - cur], summary, nvar,
                 fstride(), cap32, stream);
        // the new table: at most 2^R leaves per old leaf, and no more than the capacity
        const uint64_t most  = std::min<uint64_t>(static_cast<uint64_t>(nn) << rank, cap);
        const uint64_t items = most * static_cast<uint64_t>(ndir);
        if (rank == 2)
            hipLaunchKernelGGL(k_build_table<2>, dim3(grid1d(items, 256)), dim3(256), 0, stream, keys[1 - kc],
                               static_cast<uint32_t>(most), depth, links, metas, summary, cap32);
        else
            hipLaunchKernelGGL(k_build_table<3>, dim3(grid1d(items, 256)), dim3(256), 0, stream, keys[1 - kc],
                               static_cast<uint32_t>(most), depth, links, metas, summary, cap32);
        check_launch("k_build_table");
        publish_wait(summary, nullptr);
        const regrid_summary s = box->summary;
        if (s.err & 2u) throw amr_error(AMR_ERR_UNSUPPORTED, "coarsening to the root level is not supported");
        if (s.n_refine == 0 && s.n_coarsen == 0) return false;
        if (s.total > cap) throw amr_error(AMR_ERR_CAPACITY, "regrid would exceed the tree capacity");
        cur = 1 - cur;
        kc  = 1 - kc;
        n   = s.total;
        runs.valid = false;
        if (s.err & 1u)
        {
            // the table may hold -1 entries: no kernel may run on it again
            broken = true;
            throw amr_error(AMR_ERR_TOPOLOGY, "leaves do not tile the domain or violate 2:1 face balance");
        }
        max_level = static_cast<int>(s.max_level);
        return true;
    }
};

// A batch of independent trees with the same patch shape, concatenated into
// one leaf list so that one launch steps all of them: each instance keeps its
// own velocity, cfl and dt (through the per-instance constant table).
struct amr_batch
{
    ops                      op_table{};
    ops const*               op = &op_table;
    hipStream_t              stream{};
    int                      ni = 0, ndir = 0;
    uint64_t                 total = 0;
    std::vector<amr_ctx*>    src;
    std::vector<uint64_t>    off;
    std::vector<step_params> host_table;
    double*                  u[2]{};
    int                      cur     = 0;
    int                      variant = 0;
    uint64_t                 sweeps  = 0;
    uint64_t*                keys    = nullptr;
    face_link*               links   = nullptr;
    uint32_t*                metas   = nullptr;
    uint16_t*                inst    = nullptr;
    step_params*             table   = nullptr;
    run_lists                runs;
    struct graph_entry
    {
        uint64_t        steps;
        int             parity;
        int             variant;
        hipGraphExec_t  exec;
    };
    std::vector<graph_entry> graphs;

    ~amr_batch()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& g : graphs) (void)hipGraphExecDestroy(g.exec);
        for (void* p : { (void*)u[0], (void*)u[1], (void*)keys, (void*)links, (void*)metas, (void*)inst, (void*)table })
            if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void launch(int v)
    {
        step_params prm = host_table[0];
        prm.reverse     = static_cast<uint32_t>(sweeps++ & 1);
        op->step(*op, v, u[cur], u[1 - cur], keys, links, metas, prm, table, inst, static_cast<uint32_t>(total), runs.view(),
                 stream);
        cur = 1 - cur;
    }

    // Every instance advances `steps` steps with its own dt.  With use_graph
    // the launches are captured once per (steps, buffer parity, variant) in a
    // hipGraph and replayed as one graph launch.
    void advance(uint64_t steps, bool use_graph)
    {
        if (!use_graph)
        {
            for (uint64_t s = 0; s < steps; ++s) launch(variant);
            return;
        }
        hipGraphExec_t exec = nullptr;
        for (auto const& g : graphs)
            if (g.steps == steps && g.parity == cur && g.variant == variant) exec = g.exec;
        const int start = cur;
        if (!exec)
        {
            hipGraph_t graph = nullptr;
            AMR_HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            try
            {
                for (uint64_t s = 0; s < steps; ++s) launch(variant);
            }
            catch (...)
            {
                (void)hipStreamEndCapture(stream, &graph);
                if (graph) (void)hipGraphDestroy(graph);
                cur = start;
                throw;
            }
            AMR_HIP_CHECK(hipStreamEndCapture(stream, &graph));
            const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) throw amr_error(AMR_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
            graphs.push_back({ steps, start, variant, exec });
            cur = start;
        }
        AMR_HIP_CHECK(hipGraphLaunch(exec, stream));
        cur = (steps % 2) ? 1 - start : start;
    }
};

// Refine flags of the criterion on the current state, into `out` (one per
// leaf; ctx->flags unless given).
static void run_criterion(amr_ctx* ctx, amr_criterion const* crit, int8_t* out = nullptr)
{
    ctx->ensure_usable();
    if (!crit) throw amr_error(AMR_ERR_ARGUMENT, "null criterion");
    if (crit->kind != AMR_CRITERION_MAX_VALUE && crit->kind != AMR_CRITERION_GRADIENT)
        throw amr_error(AMR_ERR_ARGUMENT, "unknown criterion kind");
    ctx->op->criterion(*ctx->op, crit->kind, ctx->u[ctx->cur], ctx->keys[ctx->kc], ctx->links, ctx->metas,
                       crit->refine_threshold, crit->coarsen_threshold, crit->min_level, crit->max_level,
                       out ? out : ctx->flags, static_cast<uint32_t>(ctx->n), ctx->stream);
}

// Staging of amr_tree_compute_flags_noalloc: one flag per leaf, for trees of
// up to kNoAllocMaxLeaves leaves (1 MiB).  Every GPU the library runs on
// holds its own copy until the process ends.  Host code uses the address
// hipGetSymbolAddress gives on each device (noalloc_staging), never the
// array's name, which on the host is only a stand-in for the symbol.
constexpr uint64_t kNoAllocMaxLeaves = uint64_t{ 1 } << 20;
static __device__ __align__(256) int8_t scratch[kNoAllocMaxLeaves];

namespace
{
constexpr int kMaxDevices = 64;
std::mutex    g_scratch_mutex;
int8_t*       g_scratch_address[kMaxDevices] = {};

// The device address of `scratch` on `device`, looked up once per device
// with that device current (the caller's current device is restored).
int8_t* noalloc_staging(int device)
{
    if (device < 0 || device >= kMaxDevices) throw amr_error(AMR_ERR_UNSUPPORTED, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_address[device])
    {
        int prev = 0;
        AMR_HIP_CHECK(hipGetDevice(&prev));
        if (prev != device) AMR_HIP_CHECK(hipSetDevice(device));
        void*            p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (prev != device) (void)hipSetDevice(prev);
        if (e != hipSuccess)
            throw amr_error(AMR_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_address[device] = static_cast<int8_t*>(p);
    }
    return g_scratch_address[device];
}
} // namespace

// The recorded forward pass of a solve with regrids, for the discrete
// adjoint: every topology the solve went through, the dt of every step, the
// refine flags and old->new maps of every regrid, and the state entering
// every `ckpt_every`-th step.  Independent of the context once recorded.
// forward() replays the recorded topology and dt sequence from a given
// initial state and velocity (refreshing the checkpoints); adjoint()
// differentiates the most recent forward pass (the recording or a replay).
struct amr_tape
{
    struct topology
    {
        uint64_t   n     = 0;
        uint64_t*  keys  = nullptr;
        face_link* links = nullptr;
        uint32_t*  metas = nullptr;
    };
    struct regrid_record
    {
        uint64_t            step    = 0;
        bool                changed = false;
        uint64_t            n_old = 0, n_new = 0;
        int                 topo_before = 0;
        std::vector<int8_t> flags;
        uint8_t*            mark   = nullptr;
        uint8_t*            head   = nullptr;
        uint32_t*           counts = nullptr;
        uint32_t*           pos    = nullptr;
    };

    ops                        op_table{};
    ops const*                 op = &op_table;
    amr_tree_config            cfg{};
    int                        rank = 0, depth = 0, ni = 0, ndir = 0, variant = 0;
    hipStream_t                stream{};
    uint64_t                   steps = 0, ckpt_every = 1;
    double                     vel[3]{};
    std::vector<topology>      topos;
    std::vector<int>           topo_of_step;   // topology index of step k
    std::vector<int>           regrid_at_step; // regrid record before step k, or -1
    std::vector<double>        dts;
    std::vector<regrid_record> regrids;
    std::vector<double*>       ckpts; // state entering step s * ckpt_every
    uint64_t                   max_n = 0;
    // work buffers (max_n leaves each)
    double*                    lam[2]{};
    double*                    work  = nullptr;
    double*                    vpart       = nullptr; // per-leaf v-gradient triples of one segment's steps
    double*                    step_sums   = nullptr; // the triple sums of every step, step after step
    uint32_t*                  step_leaves = nullptr; // leaf count of every step
    double*                    chunk     = nullptr;
    double*                    acc       = nullptr;
    // Per-step constants in device memory, written before every pass: the
    // velocity-dependent ones change between calls while the captured
    // launches stay the same.  inst0: instance 0 for every leaf (the step
    // kernels' per-leaf table lookup).
    step_params*               step_tab = nullptr;
    adjoint_params*            adj_tab  = nullptr;
    uint16_t*                  inst0    = nullptr;
    std::vector<step_params>   h_step_tab;
    std::vector<adjoint_params> h_adj_tab;
    std::vector<uint32_t>      h_leaves;
    uint64_t*                  scratch_keys    = nullptr;
    regrid_summary*            scratch_summary = nullptr;
    std::vector<double*>       seg; // recomputed states of one segment
    // The launch sequence of a recorded tape is fixed (steps, regrids,
    // buffers), so the forward replay and the adjoint pass are each captured
    // once into a hipGraph and replayed; the kernels read the velocity-
    // dependent constants from step_tab / adj_tab, so a new velocity needs no
    // new capture.
    struct pass_graph
    {
        hipGraphExec_t exec   = nullptr;
        double*        result = nullptr; // buffer that holds the pass's output
    };
    pass_graph g_forward, g_adjoint;

    ~amr_tape()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        for (pass_graph* g : { &g_forward, &g_adjoint })
            if (g->exec) (void)hipGraphExecDestroy(g->exec);
        std::vector<void*> ptrs = { lam[0], lam[1], work, vpart, step_sums, step_leaves, chunk, acc, scratch_keys,
                                    scratch_summary, step_tab, adj_tab, inst0 };
        for (auto& t : topos) ptrs.insert(ptrs.end(), { t.keys, t.links, t.metas });
        for (auto& r : regrids) ptrs.insert(ptrs.end(), { r.mark, r.head, r.counts, r.pos });
        for (double* c : ckpts) ptrs.push_back(c);
        for (double* s : seg) ptrs.push_back(s);
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }

    double cell_size(int level, int axis) const
    {
        const double pm   = static_cast<double>(1u << (depth - level));
        const double maxc = static_cast<double>(1u << depth);
        return (cfg.domain_length[axis] * pm / maxc) / static_cast<double>(op->ext[cfg.rank - 1 - axis]);
    }

    step_params step_consts(double dt) const
    {
        step_params p{};
        for (int d = 0; d < 3; ++d)
        {
            p.vel[d] = vel[d];
            p.hs[d]  = 0.5 * std::abs(vel[d]);
        }
        for (int l = 0; l <= depth; ++l)
            for (int d = 0; d < rank; ++d) p.dtdx[l][d] = dt / cell_size(l, d);
        return p;
    }

    adjoint_params adjoint_consts(double dt) const
    {
        adjoint_params p{};
        for (int d = 0; d < 3; ++d)
        {
            p.vp[d]  = 0.5 * (vel[d] + std::abs(vel[d]));
            p.vm[d]  = 0.5 * (vel[d] - std::abs(vel[d]));
            p.sgn[d] = vel[d] > 0.0 ? 1.0 : (vel[d] < 0.0 ? -1.0 : 0.0);
        }
        for (int l = 0; l <= depth; ++l)
            for (int d = 0; d < rank; ++d) p.dtdx[l][d] = dt / cell_size(l, d);
        return p;
    }

    topology const& topo(uint64_t k) const { return topos[static_cast<size_t>(topo_of_step[k])]; }
    size_t          field_bytes(uint64_t leaves) const { return leaves * static_cast<size_t>(ni) * sizeof(double); }

    void snapshot_topology(amr_ctx const* c)
    {
        topology t;
        t.n     = c->n;
        t.keys  = dev_alloc<uint64_t>(c->n);
        t.links = dev_alloc<face_link>(c->n * ndir);
        t.metas = dev_alloc<uint32_t>(c->n * ndir);
        AMR_HIP_CHECK(hipMemcpyAsync(t.keys, c->keys[c->kc], c->n * sizeof(uint64_t), hipMemcpyDeviceToDevice, c->stream));
        AMR_HIP_CHECK(hipMemcpyAsync(t.links, c->links, c->n * ndir * sizeof(face_link), hipMemcpyDeviceToDevice, c->stream));
        AMR_HIP_CHECK(hipMemcpyAsync(t.metas, c->metas, c->n * ndir * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
        topos.push_back(t);
        max_n = std::max(max_n, c->n);
    }

    // Records `count` steps of the context's solve, regridding with `crit`
    // before every step k > 0 with k % regrid_every == 0.  The context ends
    // at the final state.
    void record(amr_ctx* c, const amr_criterion* crit, uint64_t count, uint64_t regrid_every)
    {
        steps = count;
        for (int d = 0; d < 3; ++d) vel[d] = c->vel[d];
        snapshot_topology(c);
        topo_of_step.assign(count, 0);
        regrid_at_step.assign(count, -1);
        dts.assign(count, 0.0);
        for (uint64_t k = 0; k < count; ++k)
        {
            if (crit && regrid_every > 0 && k > 0 && k % regrid_every == 0)
            {
                regrid_record r;
                r.step        = k;
                r.n_old       = c->n;
                r.topo_before = static_cast<int>(topos.size()) - 1;
                run_criterion(c, crit);
                r.flags.resize(c->n);
                AMR_HIP_CHECK(hipMemcpyAsync(r.flags.data(), c->flags, c->n, hipMemcpyDeviceToHost, c->stream));
                c->sync();
                r.changed = c->regrid();
                r.n_new   = c->n;
                if (r.changed)
                {
                    r.mark   = dev_alloc<uint8_t>(r.n_old);
                    r.head   = dev_alloc<uint8_t>(r.n_old);
                    r.counts = dev_alloc<uint32_t>(r.n_old);
                    r.pos    = dev_alloc<uint32_t>(r.n_old);
                    AMR_HIP_CHECK(hipMemcpyAsync(r.mark, c->mark, r.n_old, hipMemcpyDeviceToDevice, c->stream));
                    AMR_HIP_CHECK(hipMemcpyAsync(r.head, c->head, r.n_old, hipMemcpyDeviceToDevice, c->stream));
                    AMR_HIP_CHECK(hipMemcpyAsync(r.counts, c->counts, r.n_old * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
                    AMR_HIP_CHECK(hipMemcpyAsync(r.pos, c->pos, r.n_old * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
                    snapshot_topology(c);
                }
                regrid_at_step[k] = static_cast<int>(regrids.size());
                regrids.push_back(std::move(r));
            }
            topo_of_step[k] = static_cast<int>(topos.size()) - 1;
            if (k % ckpt_every == 0)
            {
                ckpts.push_back(dev_alloc<double>(c->n * ni));
                AMR_HIP_CHECK(hipMemcpyAsync(ckpts.back(), c->u[c->cur], field_bytes(c->n), hipMemcpyDeviceToDevice, c->stream));
            }
            dts[k] = c->compute_dt();
            c->step(dts[k]);
        }
        c->sync();
        // work buffers
        for (int b = 0; b < 2; ++b) lam[b] = dev_alloc<double>(max_n * ni);
        work            = dev_alloc<double>(max_n * ni);
        // The adjoint steps of a segment write their per-leaf partials side
        // by side; after the segment each step's are summed in a fixed order
        // into its slot of step_sums, and after the sweep the slots are summed
        // in step order: grad_v is the same bits for every checkpoint
        // interval, and the partials need one segment's worth of memory.
        vpart           = dev_alloc<double>(ckpt_every * max_n * 3);
        step_sums       = dev_alloc<double>(steps * 3);
        step_leaves     = dev_alloc<uint32_t>(steps);
        chunk           = dev_alloc<double>(kReduceBlocks * 3);
        acc             = dev_alloc<double>(3);
        scratch_keys    = dev_alloc<uint64_t>(max_n);
        scratch_summary = dev_alloc<regrid_summary>(1);
        step_tab        = dev_alloc<step_params>(steps);
        adj_tab         = dev_alloc<adjoint_params>(steps);
        inst0           = dev_alloc<uint16_t>(max_n);
        AMR_HIP_CHECK(hipMemsetAsync(inst0, 0, max_n * sizeof(uint16_t), stream));
        h_leaves.resize(steps);
        for (uint64_t k = 0; k < steps; ++k) h_leaves[k] = static_cast<uint32_t>(topo(k).n);
        AMR_HIP_CHECK(hipMemcpyAsync(step_leaves, h_leaves.data(), steps * sizeof(uint32_t), hipMemcpyHostToDevice,
                                     stream));
        for (uint64_t j = 1; j < ckpt_every; ++j) seg.push_back(dev_alloc<double>(max_n * ni));
        upload_constants();
    }

    // Every step's forward and adjoint constants for the current velocity,
    // copied to the device ahead of the next pass (stream order); skipped
    // when the device tables already hold this velocity's.
    double uploaded_v[3]{};
    bool   uploaded = false;
    void   upload_constants()
    {
        if (uploaded && std::memcmp(uploaded_v, vel, sizeof(vel)) == 0) return;
        AMR_HIP_CHECK(hipStreamSynchronize(stream)); // the host tables of the last upload are free
        h_step_tab.resize(steps);
        h_adj_tab.resize(steps);
        for (uint64_t k = 0; k < steps; ++k)
        {
            h_step_tab[k] = step_consts(dts[k]);
            h_adj_tab[k]  = adjoint_consts(dts[k]);
        }
        AMR_HIP_CHECK(hipMemcpyAsync(step_tab, h_step_tab.data(), steps * sizeof(step_params), hipMemcpyHostToDevice,
                                     stream));
        AMR_HIP_CHECK(hipMemcpyAsync(adj_tab, h_adj_tab.data(), steps * sizeof(adjoint_params), hipMemcpyHostToDevice,
                                     stream));
        std::memcpy(uploaded_v, vel, sizeof(vel));
        uploaded = true;
    }

    regrid_record const* regrid_before(uint64_t k) const
    {
        const int r = regrid_at_step[k];
        return (r >= 0 && regrids[static_cast<size_t>(r)].changed) ? &regrids[static_cast<size_t>(r)] : nullptr;
    }

    // The recorded regrid transfer (old -> new leaves) applied to src.
    void replay_transfer(regrid_record const& r, const double* src, double* dst)
    {
        op->emit(*op, topos[static_cast<size_t>(r.topo_before)].keys, src, static_cast<uint32_t>(r.n_old), depth, r.mark,
                 r.head, r.counts, r.pos, scratch_keys, dst, scratch_summary, 1, 0, 0u, stream);
    }

    void forward_step(uint64_t k, const double* src, double* dst)
    {
        topology const& t = topo(k);
        step_params     prm{}; // the sweep direction; the constants come from step_tab[k]
        prm.reverse = static_cast<uint32_t>(k & 1);
        op->step(*op, variant, src, dst, t.keys, t.links, t.metas, prm, step_tab + k, inst0, static_cast<uint32_t>(t.n),
                 step_aux{}, stream);
    }

    void forward(const double* u0, const double* v, double* uT)
    {
        AMR_HIP_CHECK(hipDeviceSynchronize()); // inputs written on other streams
        for (int d = 0; d < 3; ++d) vel[d] = d < rank ? v[d] : 0.0;
        upload_constants();
        AMR_HIP_CHECK(hipMemcpyAsync(lam[0], u0, field_bytes(topos[0].n), hipMemcpyDeviceToDevice, stream));
        run_pass(g_forward, [this]() {
            double* a = lam[0];
            double* b = lam[1];
            for (uint64_t k = 0; k < steps; ++k)
            {
                if (auto const* r = regrid_before(k))
                {
                    replay_transfer(*r, a, b);
                    std::swap(a, b);
                }
                if (k % ckpt_every == 0)
                    AMR_HIP_CHECK(hipMemcpyAsync(ckpts[k / ckpt_every], a, field_bytes(topo(k).n),
                                                 hipMemcpyDeviceToDevice, stream));
                forward_step(k, a, b);
                std::swap(a, b);
            }
            return a;
        });
        AMR_HIP_CHECK(hipMemcpyAsync(uT, g_forward.result, field_bytes(topo(steps - 1).n), hipMemcpyDeviceToDevice,
                                     stream));
        AMR_HIP_CHECK(hipStreamSynchronize(stream));
    }

    // Gradient of lamT . u(T) with respect to u0 and v (dt held fixed),
    // recomputing the states inside each checkpoint segment.
    void adjoint(const double* lamT, double* grad_u0, double* grad_v)
    {
        AMR_HIP_CHECK(hipDeviceSynchronize());
        AMR_HIP_CHECK(hipMemcpyAsync(lam[0], lamT, field_bytes(topo(steps - 1).n), hipMemcpyDeviceToDevice, stream));
        run_pass(g_adjoint, [this]() { return adjoint_sweep(); });
        AMR_HIP_CHECK(hipMemcpyAsync(grad_u0, g_adjoint.result, field_bytes(topos[0].n), hipMemcpyDeviceToDevice, stream));
        double h[3];
        AMR_HIP_CHECK(hipMemcpyAsync(h, acc, sizeof(h), hipMemcpyDeviceToHost, stream));
        AMR_HIP_CHECK(hipStreamSynchronize(stream));
        for (int d = 0; d < 3; ++d) grad_v[d] = d < rank ? h[d] : 0.0;
    }

    // The backward sweep from lambda_T in lam[0]: grad_u0 in the returned
    // buffer, grad_v in acc.
    double* adjoint_sweep()
    {
        double*                    l  = lam[0];
        double*                    lo = lam[1];
        std::vector<const double*> states(ckpt_every);
        const uint64_t             nseg = (steps + ckpt_every - 1) / ckpt_every;
        for (uint64_t s = nseg; s-- > 0;)
        {
            const uint64_t c = s * ckpt_every, e = std::min(steps, c + ckpt_every);
            states[0]        = ckpts[s];
            for (uint64_t k = c; k + 1 < e; ++k)
            {
                double* next = seg[k - c];
                if (auto const* r = regrid_before(k + 1))
                {
                    forward_step(k, states[k - c], work);
                    replay_transfer(*r, work, next);
                }
                else
                    forward_step(k, states[k - c], next);
                states[k + 1 - c] = next;
            }
            for (uint64_t k = e; k-- > c;)
            {
                topology const& t = topo(k);
                op->adj_step(*op, l, states[k - c], t.keys, t.links, t.metas, adj_tab + k, lo,
                             vpart + (k - c) * max_n * 3, static_cast<uint32_t>(t.n), stream);
                std::swap(l, lo);
                if (auto const* r = regrid_before(k))
                {
                    op->adj_regrid(*op, l, r->mark, r->head, r->counts, r->pos, lo, static_cast<uint32_t>(r->n_old), stream);
                    std::swap(l, lo);
                }
            }
            // the segment's per-step velocity partials, each step into its slot
            hipLaunchKernelGGL(k_sum3_steps, dim3(static_cast<uint32_t>(e - c)), dim3(256), 0, stream, vpart,
                               static_cast<uint64_t>(max_n), step_leaves, c, step_sums);
            check_launch("k_sum3_steps");
        }
        hipLaunchKernelGGL(k_sum3_chunks, dim3(kReduceBlocks), dim3(256), 0, stream, step_sums, steps, chunk);
        check_launch("k_sum3_chunks");
        hipLaunchKernelGGL(k_sum3_final, dim3(1), dim3(256), 0, stream, chunk, acc);
        check_launch("k_sum3_final");
        return l;
    }

    // Replays g, or captures body() into it first.  body returns the buffer
    // that holds the output.
    template <class F>
    void run_pass(pass_graph& g, F&& body)
    {
        if (g.exec)
        {
            AMR_HIP_CHECK(hipGraphLaunch(g.exec, stream));
            return;
        }
        hipGraph_t graph = nullptr;
        AMR_HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        try
        {
            g.result = body();
        }
        catch (...)
        {
            (void)hipStreamEndCapture(stream, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        AMR_HIP_CHECK(hipStreamEndCapture(stream, &graph));
        const hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess)
        {
            g.exec = nullptr;
            throw amr_error(AMR_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
        AMR_HIP_CHECK(hipGraphLaunch(g.exec, stream));
    }
};

extern "C"
{

const char* amr_last_error(void) { return g_err.c_str(); }

// The interior extents of the configured patch (layout dims 0..rank-1),
// checked: static_shape<N, M[, O]> (include/containers/static_layout.hpp)
// with every extent even (the children split each dim in half) and at
// least twice the halo width (a ghost band lies within one finer cell).
static void patch_extents(const amr_tree_config* cfg, int32_t* ext)
{
    const bool given = cfg->patch_extents[0] || cfg->patch_extents[1] || cfg->patch_extents[2];
    if (cfg->halo_width < 1 || cfg->halo_width > 8) throw amr_error(AMR_ERR_ARGUMENT, "halo_width must be 1..8");
    int64_t cells = 1;
    for (int d = 0; d < cfg->rank; ++d)
    {
        ext[d] = given ? cfg->patch_extents[d] : cfg->patch_size;
        if (given && cfg->patch_size != 0 && cfg->patch_size != ext[d])
            throw amr_error(AMR_ERR_ARGUMENT, "patch_size and patch_extents disagree (leave patch_size 0)");
        if (ext[d] < 2 || ext[d] > 4096 || ext[d] % 2 != 0)
            throw amr_error(AMR_ERR_ARGUMENT, "patch extents must be even and within 2..4096");
        if (ext[d] < 2 * cfg->halo_width)
            throw amr_error(AMR_ERR_ARGUMENT, "patch extents must be at least twice the halo width");
        cells *= ext[d] + 2 * cfg->halo_width;
    }
    if (given && cfg->rank == 2 && cfg->patch_extents[2] != 0 && cfg->patch_extents[2] != 1)
        throw amr_error(AMR_ERR_ARGUMENT, "a 2D patch has no third extent");
    if (cells > (int64_t{ 1 } << 24)) throw amr_error(AMR_ERR_ARGUMENT, "padded patch larger than 2^24 cells");
}

// ndtree(capacity) with `nvar` fields per cell (ndtree.hpp:320-425).
static std::unique_ptr<amr_ctx> create_ctx(const amr_tree_config* cfg, int nvar)
{
    if (cfg->rank != 2 && cfg->rank != 3) throw amr_error(AMR_ERR_ARGUMENT, "rank must be 2 or 3");
    if (cfg->max_depth < 0 || cfg->max_depth > kMaxDepth || cfg->rank * cfg->max_depth + 6 > 64)
        throw amr_error(AMR_ERR_ARGUMENT, "max_depth out of range");
    if (cfg->capacity < 1 || cfg->capacity > 0xFFFFFFFFull / 8)
        throw amr_error(AMR_ERR_ARGUMENT, "capacity out of range");
    int32_t ext[3] = { 1, 1, 1 };
    patch_extents(cfg, ext);
    auto c      = std::make_unique<amr_ctx>();
    c->cfg      = *cfg;
    c->op_table = shape_ops(cfg->rank, ext, cfg->halo_width);
    c->rank     = cfg->rank;
    c->depth    = cfg->max_depth;
    c->ndir     = 2 * cfg->rank;
    c->cap      = cfg->capacity;
    c->nvar     = nvar;
    c->ni       = c->op->ni();
    c->flat     = c->op->padded_cells();
    AMR_HIP_CHECK(hipGetDevice(&c->device));
    AMR_HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    const size_t cap = cfg->capacity;
    for (int b = 0; b < 2; ++b)
    {
        c->u[b]    = dev_alloc<double>(cap * c->ni * nvar);
        c->keys[b] = dev_alloc<uint64_t>(cap);
        AMR_HIP_CHECK(hipMemsetAsync(c->u[b], 0, cap * c->ni * nvar * sizeof(double), c->stream));
    }
    c->links      = dev_alloc<face_link>(cap * c->ndir);
    c->metas      = dev_alloc<uint32_t>(cap * c->ndir);
    c->flags      = dev_alloc<int8_t>(cap);
    c->mark       = dev_alloc<uint8_t>(cap);
    c->head       = dev_alloc<uint8_t>(cap);
    c->counts     = dev_alloc<uint32_t>(cap);
    c->pos        = dev_alloc<uint32_t>(cap);
    c->block_sums = dev_alloc<uint32_t>(cap / kScanBlock + 2);
    c->summary    = dev_alloc<regrid_summary>(1);
    AMR_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&c->box), sizeof(host_mailbox), 0));
    std::memset(c->box, 0, sizeof(host_mailbox));
    // the periodic root (ndtree.hpp:411-421)
    const uint64_t root = 0;
    AMR_HIP_CHECK(hipMemcpyAsync(c->keys[0], &root, sizeof(root), hipMemcpyHostToDevice, c->stream));
    c->n         = 1;
    c->max_level = 0;
    if (run_lists::wanted(c->op)) c->runs.allocate(cap);
    c->build_table();
    return c;
}

int amr_tree_create(const amr_tree_config* cfg, amr_ctx** out)
{
    return guarded(
        [&]
        {
            if (!cfg || !out) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            *out = nullptr;
            *out = create_ctx(cfg, 1).release();
        }
    );
}

int amr_euler_create(const amr_tree_config* cfg, double gamma, double cfl, amr_ctx** out)
{
    return guarded(
        [&]
        {
            if (!cfg || !out) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            *out = nullptr;
            if (!(gamma > 1.0) || !(cfl > 0.0)) throw amr_error(AMR_ERR_ARGUMENT, "need gamma > 1 and cfl > 0");
            auto c   = create_ctx(cfg, cfg->rank + 2);
            c->gamma = gamma;
            c->cfl   = cfl;
            c->esc   = dev_alloc<euler_scalars>(1);
            *out = c.release();
        }
    );
}

int amr_tree_destroy(amr_ctx* ctx)
{
    return guarded([&] { delete ctx; });
}

int amr_tree_size(const amr_ctx* ctx, uint64_t* n)
{
    return guarded(
        [&]
        {
            if (!ctx || !n) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            *n = ctx->n;
        }
    );
}

int amr_tree_get_keys(amr_ctx* ctx, uint64_t* host_keys, uint64_t n)
{
    return guarded(
        [&]
        {
            if (!ctx || (!host_keys && n)) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            if (n < ctx->n) throw amr_error(AMR_ERR_ARGUMENT, "key buffer smaller than the leaf count");
            AMR_HIP_CHECK(hipMemcpyAsync(host_keys, ctx->keys[ctx->kc], ctx->n * sizeof(uint64_t),
                                         hipMemcpyDeviceToHost, ctx->stream));
            ctx->sync();
        }
    );
}

int amr_tree_get_neighbor_metadata(amr_ctx* ctx, amr_halo_direction_metadata* host, uint64_t count)
{
    return guarded(
        [&]
        {
            if (!ctx || (!host && count)) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            const uint64_t m = ctx->n * ctx->ndir;
            if (count < m) throw amr_error(AMR_ERR_ARGUMENT, "metadata buffer too small");
            std::vector<face_link> links(m);
            std::vector<uint32_t>  metas(m);
            AMR_HIP_CHECK(hipMemcpyAsync(links.data(), ctx->links, m * sizeof(face_link), hipMemcpyDeviceToHost,
                                         ctx->stream));
            AMR_HIP_CHECK(hipMemcpyAsync(metas.data(), ctx->metas, m * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                         ctx->stream));
            ctx->sync();
            const int k = 1 << (ctx->rank - 1);
            for (uint64_t e = 0; e < m; ++e)
            {
                amr_halo_direction_metadata r{};
                r.neighbor = -1;
                for (int j = 0; j < 4; ++j) r.finer_ids[j] = -1;
                const uint32_t rel = rel_of(metas[e]);
                r.relation         = static_cast<int8_t>(rel);
                if (rel == REL_SAME) r.neighbor = links[e].ids[0];
                if (rel == REL_FINER)
                    for (int j = 0; j < k; ++j) r.finer_ids[j] = links[e].ids[j];
                if (rel == REL_COARSER)
                {
                    r.neighbor = links[e].ids[0];
                    for (int d = 0; d < ctx->rank; ++d) r.quadrant[d] = static_cast<int32_t>(quad_of(metas[e], d));
                }
                host[e] = r;
            }
        }
    );
}

int amr_tree_compute_flags(amr_ctx* ctx, const amr_criterion* crit, int8_t* host_flags, uint64_t n)
{
    return guarded(
        [&]
        {
            if (!ctx || (!host_flags && n)) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            if (n < ctx->n) throw amr_error(AMR_ERR_ARGUMENT, "flag buffer smaller than the leaf count");
            run_criterion(ctx, crit);
            AMR_HIP_CHECK(hipMemcpyAsync(host_flags, ctx->flags, ctx->n, hipMemcpyDeviceToHost, ctx->stream));
            ctx->sync();
        }
    );
}

int amr_tree_compute_flags_noalloc(amr_ctx* ctx, const amr_criterion* crit, int8_t* host_flags, uint64_t n)
{
    return guarded(
        [&]
        {
            if (!ctx || (!host_flags && n)) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            if (n < ctx->n) throw amr_error(AMR_ERR_ARGUMENT, "flag buffer smaller than the leaf count");
            if (ctx->n > kNoAllocMaxLeaves)
                throw amr_error(AMR_ERR_CAPACITY, "amr_tree_compute_flags_noalloc: " + std::to_string(ctx->n) +
                                                      " leaves exceed its staging of " +
                                                      std::to_string(kNoAllocMaxLeaves) + " leaves");
            int8_t* stage = noalloc_staging(ctx->device);
            // every leaf's flag is written before the copy: nothing in the
            // staging needs clearing
            run_criterion(ctx, crit, stage);
            AMR_HIP_CHECK(hipMemcpyAsync(host_flags, stage, ctx->n, hipMemcpyDeviceToHost, ctx->stream));
            ctx->sync();
        }
    );
}

int amr_tree_reconstruct_flags(amr_ctx* ctx, const int8_t* host_flags, uint64_t n, int* changed)
{
    return guarded(
        [&]
        {
            if (!ctx || (!host_flags && n)) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            if (n != ctx->n) throw amr_error(AMR_ERR_ARGUMENT, "one flag per leaf is required");
            AMR_HIP_CHECK(hipMemcpyAsync(ctx->flags, host_flags, n, hipMemcpyHostToDevice, ctx->stream));
            const bool c = ctx->regrid();
            if (changed) *changed = c ? 1 : 0;
        }
    );
}

int amr_tree_reconstruct_criterion(amr_ctx* ctx, const amr_criterion* crit, int* changed)
{
    return guarded(
        [&]
        {
            if (!ctx) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            run_criterion(ctx, crit);
            const bool c = ctx->regrid();
            if (changed) *changed = c ? 1 : 0;
        }
    );
}

int amr_tree_set_interior(amr_ctx* ctx, const double* host, uint64_t count)
{
    return guarded(
        [&]
        {
            if (!ctx || (!host && count)) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            const uint64_t per = ctx->n * ctx->ni;
            if (count != per * ctx->nvar) throw amr_error(AMR_ERR_ARGUMENT, "count must be fields * size * n^rank");
            for (int v = 0; v < ctx->nvar; ++v)
                AMR_HIP_CHECK(hipMemcpyAsync(ctx->u[ctx->cur] + v * ctx->fstride(), host + v * per, per * sizeof(double),
                                             hipMemcpyHostToDevice, ctx->stream));
            ctx->sync();
        }
    );
}

int amr_tree_get_interior(amr_ctx* ctx, double* host, uint64_t count)
{
    return guarded(
        [&]
        {
            if (!ctx || (!host && count)) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            const uint64_t per = ctx->n * ctx->ni;
            if (count < per * ctx->nvar) throw amr_error(AMR_ERR_ARGUMENT, "buffer smaller than fields * size * n^rank");
            for (int v = 0; v < ctx->nvar; ++v)
                AMR_HIP_CHECK(hipMemcpyAsync(host + v * per, ctx->u[ctx->cur] + v * ctx->fstride(), per * sizeof(double),
                                             hipMemcpyDeviceToHost, ctx->stream));
            ctx->sync();
        }
    );
}

static void materialize_padded(amr_ctx* ctx)
{
    ctx->ensure_usable();
    if (!ctx->padded) ctx->padded = dev_alloc<double>(ctx->cap * ctx->flat * ctx->nvar);
    for (int v = 0; v < ctx->nvar; ++v)
        ctx->op->padded(*ctx->op, ctx->u[ctx->cur] + v * ctx->fstride(), ctx->padded + v * ctx->cap * ctx->flat, ctx->links,
                        ctx->metas, static_cast<uint32_t>(ctx->n), ctx->stream);
}

int amr_tree_get_patches(amr_ctx* ctx, double* host, uint64_t count)
{
    return guarded(
        [&]
        {
            if (!ctx || (!host && count)) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            const uint64_t per = ctx->n * ctx->flat;
            if (count < per * ctx->nvar)
                throw amr_error(AMR_ERR_ARGUMENT, "buffer smaller than fields * size * (n+2h)^rank");
            materialize_padded(ctx);
            for (int v = 0; v < ctx->nvar; ++v)
                AMR_HIP_CHECK(hipMemcpyAsync(host + v * per, ctx->padded + v * ctx->cap * ctx->flat, per * sizeof(double),
                                             hipMemcpyDeviceToHost, ctx->stream));
            ctx->sync();
        }
    );
}

int amr_tree_halo_exchange_update(amr_ctx* ctx)
{
    return guarded(
        [&]
        {
            if (!ctx) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            materialize_padded(ctx);
        }
    );
}

int amr_tree_swap_buffers(amr_ctx* ctx)
{
    return guarded(
        [&]
        {
            if (!ctx) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            ctx->cur       = 1 - ctx->cur;
        }
    );
}

int amr_tree_device_interior(amr_ctx* ctx, double** dev_ptr)
{
    return guarded(
        [&]
        {
            if (!ctx || !dev_ptr) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            *dev_ptr = ctx->u[ctx->cur];
            ctx->sync(); // queued steps have written it
        }
    );
}

// ---- legacy VTK output (include/ndtree/vtk_print.hpp:53-254) -------------

extern "C++"
{
namespace
{
// vtk_print::write_binary: big-endian bytes of the value converted to T.
template <typename T>
void vtk_be(std::ofstream& f, T v)
{
    unsigned char b[sizeof(T)];
    std::memcpy(b, &v, sizeof(T));
    std::reverse(b, b + sizeof(T));
    f.write(reinterpret_cast<const char*>(b), sizeof(T));
}

// Field names of the reference's cell types (include/solver/cell_types.hpp).
const char* vtk_field_name(int rank, int nvar, int v)
{
    if (nvar == 1) return "Scalar";
    static const char* const n2[] = { "Rho", "Rhou", "Rhov", "E2D" };
    static const char* const n3[] = { "Rho", "Rhou", "Rhov", "Rhow", "E3D" };
    return rank == 2 ? n2[v] : n3[v];
}
} // namespace
} // extern "C++"

int amr_tree_write_vtk(amr_ctx* ctx, const char* path)
{
    return guarded(
        [&]
        {
            if (!ctx || !path) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            ctx->ensure_usable();
            const int      R = ctx->rank, h = ctx->cfg.halo_width;
            const int32_t* ext = ctx->op->ext;
            const uint64_t leaves = ctx->n, ni = static_cast<uint64_t>(ctx->ni), cells = leaves * ni;
            std::vector<uint64_t> keys(leaves);
            std::vector<double>   u(cells * ctx->nvar);
            AMR_HIP_CHECK(hipMemcpyAsync(keys.data(), ctx->keys[ctx->kc], leaves * sizeof(uint64_t),
                                         hipMemcpyDeviceToHost, ctx->stream));
            for (int v = 0; v < ctx->nvar; ++v)
                AMR_HIP_CHECK(hipMemcpyAsync(u.data() + v * cells, ctx->u[ctx->cur] + v * ctx->fstride(),
                                             cells * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
            ctx->sync();
            std::ofstream f(path, std::ios::binary);
            if (!f) throw amr_error(AMR_ERR_ARGUMENT, std::string("cannot open ") + path);
            // vtk_print::write_header
            f << "# vtk DataFile Version 3.0\n";
            f << "AMR Tree Structure\n";
            f << "BINARY\n";
            f << "DATASET UNSTRUCTURED_GRID\n";
            const int box = 1 << R;
            f << "POINTS " << cells * box << " double\n";
            const double maxc = static_cast<double>(1u << ctx->depth);
            for (uint64_t i = 0; i < leaves; ++i)
            {
                uint32_t c[3] = { 0, 0, 0 };
                if (R == 2)
                    deinterleave<2>(keys[i] >> 6, c);
                else
                    deinterleave<3>(keys[i] >> 6, c);
                const int lv = key_level(keys[i]);
                double    org[3], cs[3];
                for (int d = 0; d < R; ++d)
                {
                    // physics_system::patch_coord / cell_sizes
                    org[d] = ctx->cfg.domain_length[d] * static_cast<double>(c[d]) / maxc;
                    cs[d]  = ctx->cell_size(lv, d);
                }
                for (uint64_t l = 0; l < ni; ++l)
                {
                    // interior cell l, row-major with x last
                    const int    ix = static_cast<int>(l % ext[R - 1]);
                    const int    iy = static_cast<int>((l / ext[R - 1]) % ext[R - 2]);
                    const double x  = org[0] + static_cast<double>(ix) * cs[0];
                    const double y  = org[1] + static_cast<double>(iy) * cs[1];
                    if (R == 2)
                    {
                        const double q[4][2] = { { x, y }, { x + cs[0], y }, { x + cs[0], y + cs[1] }, { x, y + cs[1] } };
                        for (auto const& p : q)
                        {
                            vtk_be<double>(f, p[0]);
                            vtk_be<double>(f, p[1]);
                            vtk_be<double>(f, 0.0);
                        }
                    }
                    else
                    {
                        const int    iz = static_cast<int>(l / (static_cast<uint64_t>(ext[1]) * ext[2]));
                        const double z  = org[2] + static_cast<double>(iz) * cs[2];
                        // VTK_VOXEL corner order
                        for (int k = 0; k < 8; ++k)
                        {
                            vtk_be<double>(f, (k & 1) ? x + cs[0] : x);
                            vtk_be<double>(f, (k & 2) ? y + cs[1] : y);
                            vtk_be<double>(f, (k & 4) ? z + cs[2] : z);
                        }
                    }
                }
            }
            f << "CELLS " << cells << ' ' << cells * (box + 1) << '\n';
            for (uint64_t i = 0; i < cells; ++i)
            {
                vtk_be<int>(f, box);
                for (int j = 0; j < box; ++j) vtk_be<int>(f, static_cast<int>(i * box + j));
            }
            f << "CELL_TYPES " << cells << '\n';
            for (uint64_t i = 0; i < cells; ++i) vtk_be<int>(f, R == 2 ? 9 : 11);
            f << "CELL_DATA " << cells << '\n';
            f << "SCALARS cell_index int 1\n";
            f << "LOOKUP_TABLE default\n";
            for (uint64_t i = 0; i < cells; ++i) vtk_be<int>(f, static_cast<int>(i));
            // The reference evaluates is_halo_cell on the interior element
            // index j read as a padded linear index (vtk_print.hpp:203-214);
            // the flags are reproduced as written.
            f << "SCALARS is_halo int 1\n";
            f << "LOOKUP_TABLE default\n";
            std::vector<int> halo_flag(ni);
            for (uint64_t j = 0; j < ni; ++j)
            {
                uint64_t rem = j;
                int      flag = 0;
                for (int d = 0; d < R; ++d)
                {
                    // padded layout: extent n(d) + 2h, strides of the later dims
                    uint64_t stride = 1;
                    for (int e = d + 1; e < R; ++e) stride *= static_cast<uint64_t>(ext[e] + 2 * h);
                    const uint64_t P   = static_cast<uint64_t>(ext[d] + 2 * h);
                    const uint64_t rel = rem / stride;
                    if (rel < static_cast<uint64_t>(h) || rel >= P - static_cast<uint64_t>(h)) flag = 1;
                    rem %= stride;
                    if (flag) break;
                }
                halo_flag[j] = flag;
            }
            for (uint64_t i = 0; i < leaves; ++i)
                for (uint64_t j = 0; j < ni; ++j) vtk_be<int>(f, halo_flag[j]);
            for (int v = 0; v < ctx->nvar; ++v)
            {
                f << "SCALARS " << vtk_field_name(R, ctx->nvar, v) << " double 1\n";
                f << "LOOKUP_TABLE default\n";
                for (uint64_t i = 0; i < cells; ++i) vtk_be<double>(f, u[v * cells + i]);
            }
            if (!f) throw amr_error(AMR_ERR_ARGUMENT, std::string("write failed: ") + path);
        }
    );
}

int amr_tree_stream(amr_ctx* ctx, void** hip_stream)
{
    return guarded(
        [&]
        {
            if (!ctx || !hip_stream) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            *hip_stream = reinterpret_cast<void*>(ctx->stream);
        }
    );
}

int amr_solver_configure(amr_ctx* ctx, const double velocity[3], double cfl)
{
    return guarded(
        [&]
        {
            if (!ctx || !velocity) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            if (!(cfl > 0.0)) throw amr_error(AMR_ERR_ARGUMENT, "cfl must be positive");
            if (ctx->euler()) throw amr_error(AMR_ERR_UNSUPPORTED, "an Euler context has no advection velocity");
            for (int d = 0; d < 3; ++d) ctx->vel[d] = velocity[d];
            ctx->cfl = cfl;
        }
    );
}

int amr_solver_compute_dt(amr_ctx* ctx, double* dt)
{
    return guarded(
        [&]
        {
            if (!ctx || !dt) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            *dt = ctx->euler() ? ctx->euler_dt_now() : ctx->compute_dt();
        }
    );
}

// advance_batch_async, CPU branch semantics (amr_solver.hpp:221-241): dt is
// known on the host for constant-velocity advection, so only the step
// kernels are queued.
int amr_solver_advance_batch_async(amr_ctx* ctx, uint64_t step_count, double remaining_time)
{
    return guarded(
        [&]
        {
            if (!ctx) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            ctx->pending_dt    = 0.0;
            ctx->pending_steps = 0;
            if (ctx->euler())
            {
                // dt depends on the state: the device computes it every step
                // and the host reads the sums in finish_advance_batch
                ctx->euler_advance(step_count, remaining_time);
                return;
            }
            double remaining   = remaining_time;
            uint64_t s         = 0;
            while (s < step_count)
            {
                if (remaining <= 0.0) break;
                const double dt      = ctx->compute_dt();
                const double step_dt = std::min(dt, remaining);
                if (step_dt <= 0.0) break;
                const double rest = remaining - step_dt;
                // the following step's dt, as the loop would take it
                const double next_dt = (s + 1 < step_count && rest > 0.0) ? std::min(dt, rest) : 0.0;
                if (next_dt > 0.0 && ctx->fusable())
                {
                    ctx->step2(step_dt, next_dt);
                    ctx->pending_dt += step_dt;
                    ctx->pending_dt += next_dt;
                    remaining = rest - next_dt;
                    ctx->pending_steps += 2;
                    s += 2;
                    continue;
                }
                ctx->step(step_dt);
                ctx->pending_dt += step_dt;
                remaining = rest;
                ++ctx->pending_steps;
                ++s;
            }
        }
    );
}

int amr_solver_finish_advance_batch(amr_ctx* ctx, double* dt_sum, uint64_t* executed_steps)
{
    return guarded(
        [&]
        {
            if (!ctx) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            if (ctx->euler())
            {
                const euler_scalars e = ctx->euler_read();
                ctx->pending_dt       = e.acc;
                ctx->pending_steps    = e.executed;
            }
            else
                ctx->publish_wait(nullptr, nullptr);
            if (dt_sum) *dt_sum = ctx->pending_dt;
            if (executed_steps) *executed_steps = ctx->pending_steps;
        }
    );
}

int amr_solver_advance(amr_ctx* ctx, double* dt)
{
    int rc = amr_solver_advance_batch_async(ctx, 1, std::numeric_limits<double>::max());
    if (rc != AMR_OK) return rc;
    return amr_solver_finish_advance_batch(ctx, dt, nullptr);
}

int amr_solver_step_variants(amr_ctx* ctx, int32_t* count)
{
    return guarded(
        [&]
        {
            if (!ctx || !count) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            *count = ctx->op->variant_count();
        }
    );
}

int amr_solver_step_variant_name(amr_ctx* ctx, int32_t variant, const char** name)
{
    return guarded(
        [&]
        {
            if (!ctx || !name) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            if (variant < 0 || variant >= ctx->op->variant_count()) throw amr_error(AMR_ERR_ARGUMENT, "no such variant");
            *name = ctx->op->variant_name(variant);
        }
    );
}

int amr_solver_set_step_variant(amr_ctx* ctx, int32_t variant)
{
    return guarded(
        [&]
        {
            if (!ctx) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            if (variant < 0 || variant >= ctx->op->variant_count()) throw amr_error(AMR_ERR_ARGUMENT, "no such variant");
            ctx->variant = variant;
        }
    );
}

int amr_solver_get_step_variant(amr_ctx* ctx, int32_t* variant)
{
    return guarded(
        [&]
        {
            if (!ctx || !variant) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            *variant = ctx->variant;
        }
    );
}

int amr_solver_set_step_fusion(amr_ctx* ctx, int32_t enable)
{
    return guarded(
        [&]
        {
            if (!ctx) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            ctx->fuse = enable != 0;
        }
    );
}

int amr_solver_get_step_fusion(amr_ctx* ctx, int32_t* enabled, int32_t* applies)
{
    return guarded(
        [&]
        {
            if (!ctx) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            if (enabled) *enabled = ctx->fuse ? 1 : 0;
            if (applies) *applies = ctx->fusion_applies() ? 1 : 0;
        }
    );
}

int amr_solver_autotune(amr_ctx* ctx, uint32_t trial_steps, int32_t* chosen, float* best_us)
{
    return guarded(
        [&]
        {
            if (ctx && ctx->euler()) throw amr_error(AMR_ERR_UNSUPPORTED, "the step variants are advection kernels");
            if (!ctx) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            const int v = ctx->autotune(trial_steps, best_us);
            if (chosen) *chosen = v;
        }
    );
}

int amr_tree_fields(const amr_ctx* ctx, int32_t* fields)
{
    return guarded(
        [&]
        {
            if (!ctx || !fields) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            *fields = ctx->nvar;
        }
    );
}

int amr_euler_configure(amr_ctx* ctx, double gamma, double cfl)
{
    return guarded(
        [&]
        {
            if (!ctx) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            if (!ctx->euler()) throw amr_error(AMR_ERR_UNSUPPORTED, "not an Euler context");
            if (!(gamma > 1.0) || !(cfl > 0.0)) throw amr_error(AMR_ERR_ARGUMENT, "need gamma > 1 and cfl > 0");
            ctx->gamma = gamma;
            ctx->cfl   = cfl;
        }
    );
}

int amr_euler_set_primitive(amr_ctx* ctx, const double* host_prim, uint64_t count)
{
    return guarded(
        [&]
        {
            if (!ctx || (!host_prim && count)) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            if (!ctx->euler()) throw amr_error(AMR_ERR_UNSUPPORTED, "not an Euler context");
            const uint64_t per = ctx->n * ctx->ni;
            const int      nv  = ctx->nvar;
            if (count != per * nv) throw amr_error(AMR_ERR_ARGUMENT, "count must be fields * size * n^rank");
            std::vector<double> cons(count);
            double              prim[5], out[5];
            for (uint64_t c = 0; c < per; ++c)
            {
                for (int v = 0; v < nv; ++v) prim[v] = host_prim[v * per + c];
                if (ctx->rank == 2)
                    euler_flux<2>::prim_to_cons(prim, out, ctx->gamma);
                else
                    euler_flux<3>::prim_to_cons(prim, out, ctx->gamma);
                for (int v = 0; v < nv; ++v) cons[v * per + c] = out[v];
            }
            for (int v = 0; v < nv; ++v)
                AMR_HIP_CHECK(hipMemcpyAsync(ctx->u[ctx->cur] + v * ctx->fstride(), cons.data() + v * per,
                                             per * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
            ctx->sync();
        }
    );
}

int amr_batch_create(amr_ctx* const* trees, uint32_t count, amr_batch** out)
{
    return guarded(
        [&]
        {
            for (uint32_t t = 0; trees && t < count; ++t)
                if (trees[t] && trees[t]->euler())
                    throw amr_error(AMR_ERR_UNSUPPORTED, "batches hold advection trees only");
            if (!trees || !out || count == 0) throw amr_error(AMR_ERR_ARGUMENT, "null argument or empty batch");
            if (count > 65535) throw amr_error(AMR_ERR_ARGUMENT, "at most 65535 instances per batch");
            *out = nullptr;
            auto b  = std::make_unique<amr_batch>();
            b->op_table = *trees[0]->op;
            b->ni   = trees[0]->ni;
            b->ndir = trees[0]->ndir;
            for (uint32_t i = 0; i < count; ++i)
            {
                amr_ctx* t = trees[i];
                if (!t) throw amr_error(AMR_ERR_ARGUMENT, "null tree");
                if (!t->op->same_shape(*b->op) || t->depth != trees[0]->depth)
                    throw amr_error(AMR_ERR_ARGUMENT, "batched trees must share rank, depth, patch size and halo");
                t->ensure_usable();
                t->sync();
                b->src.push_back(t);
                b->off.push_back(b->total);
                b->total += t->n;
            }
            if (b->total > 0xFFFFFFFFull / 8) throw amr_error(AMR_ERR_CAPACITY, "batch too large");
            AMR_HIP_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
            for (int k = 0; k < 2; ++k) b->u[k] = dev_alloc<double>(b->total * b->ni);
            b->keys  = dev_alloc<uint64_t>(b->total);
            b->links = dev_alloc<face_link>(b->total * b->ndir);
            b->metas = dev_alloc<uint32_t>(b->total * b->ndir);
            b->inst  = dev_alloc<uint16_t>(b->total);
            b->table = dev_alloc<step_params>(count);
            for (uint32_t i = 0; i < count; ++i)
            {
                amr_ctx const* t = b->src[i];
                const uint64_t o = b->off[i];
                AMR_HIP_CHECK(hipMemcpyAsync(b->u[0] + o * b->ni, t->u[t->cur], t->n * b->ni * sizeof(double),
                                             hipMemcpyDeviceToDevice, b->stream));
                AMR_HIP_CHECK(hipMemcpyAsync(b->keys + o, t->keys[t->kc], t->n * sizeof(uint64_t),
                                             hipMemcpyDeviceToDevice, b->stream));
                AMR_HIP_CHECK(hipMemcpyAsync(b->links + o * b->ndir, t->links, t->n * b->ndir * sizeof(face_link),
                                             hipMemcpyDeviceToDevice, b->stream));
                AMR_HIP_CHECK(hipMemcpyAsync(b->metas + o * b->ndir, t->metas, t->n * b->ndir * sizeof(uint32_t),
                                             hipMemcpyDeviceToDevice, b->stream));
                const uint64_t nl = t->n * b->ndir;
                hipLaunchKernelGGL(k_offset_links, dim3(grid1d(nl, 256)), dim3(256), 0, b->stream, b->links + o * b->ndir,
                                   nl, static_cast<int32_t>(o));
                hipLaunchKernelGGL(k_fill_instance, dim3(grid1d(t->n, 256)), dim3(256), 0, b->stream, b->inst + o, t->n,
                                   static_cast<uint16_t>(i));
                check_launch("batch packing");
                b->host_table.push_back(t->make_params(t->compute_dt()));
            }
            AMR_HIP_CHECK(hipMemcpyAsync(b->table, b->host_table.data(), count * sizeof(step_params),
                                         hipMemcpyHostToDevice, b->stream));
            if (run_lists::wanted(b->op))
            {
                b->runs.allocate(b->total);
                b->runs.build(b->op->rank, b->keys, static_cast<uint32_t>(b->total), trees[0]->depth, b->links,
                              b->metas, b->stream);
            }
            AMR_HIP_CHECK(hipStreamSynchronize(b->stream));
            *out = b.release();
        }
    );
}

int amr_batch_destroy(amr_batch* b)
{
    return guarded([&] { delete b; });
}

int amr_batch_size(amr_batch* b, uint64_t* total_leaves)
{
    return guarded(
        [&]
        {
            if (!b || !total_leaves) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            *total_leaves = b->total;
        }
    );
}

int amr_batch_advance(amr_batch* b, uint64_t steps, int use_graph)
{
    return guarded(
        [&]
        {
            if (!b) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            if (steps) b->advance(steps, use_graph != 0);
        }
    );
}

int amr_batch_synchronize(amr_batch* b)
{
    return guarded(
        [&]
        {
            if (!b) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            AMR_HIP_CHECK(hipStreamSynchronize(b->stream));
        }
    );
}

int amr_batch_scatter(amr_batch* b)
{
    return guarded(
        [&]
        {
            if (!b) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            for (size_t i = 0; i < b->src.size(); ++i)
            {
                amr_ctx* t = b->src[i];
                AMR_HIP_CHECK(hipMemcpyAsync(t->u[t->cur], b->u[b->cur] + b->off[i] * b->ni, t->n * b->ni * sizeof(double),
                                             hipMemcpyDeviceToDevice, b->stream));
            }
            AMR_HIP_CHECK(hipStreamSynchronize(b->stream));
        }
    );
}

int amr_batch_stream(amr_batch* b, void** hip_stream)
{
    return guarded(
        [&]
        {
            if (!b || !hip_stream) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            *hip_stream = reinterpret_cast<void*>(b->stream);
        }
    );
}

int amr_batch_set_step_variant(amr_batch* b, int32_t variant)
{
    return guarded(
        [&]
        {
            if (!b) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            if (variant < 0 || variant >= b->op->variant_count()) throw amr_error(AMR_ERR_ARGUMENT, "no such variant");
            b->variant = variant;
        }
    );
}

int amr_batch_autotune(amr_batch* b, uint32_t trial_steps, int32_t* chosen, float* best_us)
{
    return guarded(
        [&]
        {
            if (!b) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            step_params prm = b->host_table[0];
            b->variant      = autotune_steps(
                b->op, b->u[b->cur], b->u[1 - b->cur], b->total, b->ni,
                [&](int v, const double* src, double* dst)
                {
                    prm.reverse ^= 1u;
                    b->op->step(*b->op, v, src, dst, b->keys, b->links, b->metas, prm, b->table, b->inst,
                                static_cast<uint32_t>(b->total), b->runs.view(), b->stream);
                },
                b->stream, trial_steps, best_us);
            if (chosen) *chosen = b->variant;
        }
    );
}

int amr_ctx_synchronize(amr_ctx* ctx)
{
    return guarded(
        [&]
        {
            if (!ctx) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            ctx->sync();
        }
    );
}

int amr_tape_record(amr_ctx* ctx, const amr_criterion* criterion, uint64_t steps, uint64_t regrid_every,
                    uint64_t checkpoint_every, amr_tape** out)
{
    return guarded(
        [&]
        {
            if (ctx && ctx->euler()) throw amr_error(AMR_ERR_UNSUPPORTED, "the adjoint tape records advection solves");
            if (!ctx || !out) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            *out = nullptr;
            ctx->ensure_usable();
            if (steps == 0) throw amr_error(AMR_ERR_ARGUMENT, "at least one step is required");
            if (checkpoint_every == 0) throw amr_error(AMR_ERR_ARGUMENT, "checkpoint_every must be positive");
            if (ctx->n == 0) throw amr_error(AMR_ERR_ARGUMENT, "empty tree");
            auto t        = std::make_unique<amr_tape>();
            t->op_table   = *ctx->op;
            t->cfg        = ctx->cfg;
            t->rank       = ctx->rank;
            t->depth      = ctx->depth;
            t->ni         = ctx->ni;
            t->ndir       = ctx->ndir;
            t->variant    = ctx->variant;
            t->ckpt_every = std::min(checkpoint_every, steps);
            AMR_HIP_CHECK(hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking));
            t->record(ctx, criterion, steps, regrid_every);
            *out = t.release();
        }
    );
}

int amr_tape_destroy(amr_tape* tape)
{
    return guarded([&] { delete tape; });
}

int amr_tape_get_info(const amr_tape* tape, amr_tape_info* info)
{
    return guarded(
        [&]
        {
            if (!tape || !info) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            info->steps            = tape->steps;
            info->checkpoint_every = tape->ckpt_every;
            info->checkpoints      = tape->ckpts.size();
            info->regrids          = tape->regrids.size();
            info->topologies       = tape->topos.size();
            info->initial_leaves   = tape->topos.front().n;
            info->final_leaves     = tape->topo(tape->steps - 1).n;
            info->max_leaves       = tape->max_n;
            info->cells_per_leaf   = static_cast<uint64_t>(tape->ni);
        }
    );
}

int amr_tape_get_dts(const amr_tape* tape, double* dts, uint64_t count)
{
    return guarded(
        [&]
        {
            if (!tape || !dts) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            if (count < tape->steps) throw amr_error(AMR_ERR_ARGUMENT, "buffer smaller than the step count");
            std::copy(tape->dts.begin(), tape->dts.end(), dts);
        }
    );
}

int amr_tape_get_regrid(const amr_tape* tape, uint64_t index, uint64_t* step, uint64_t* leaves_before, int* changed,
                        int8_t* flags, uint64_t flags_count)
{
    return guarded(
        [&]
        {
            if (!tape) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            if (index >= tape->regrids.size()) throw amr_error(AMR_ERR_ARGUMENT, "regrid index out of range");
            auto const& r = tape->regrids[index];
            if (step) *step = r.step;
            if (leaves_before) *leaves_before = r.n_old;
            if (changed) *changed = r.changed ? 1 : 0;
            if (flags)
            {
                if (flags_count < r.n_old) throw amr_error(AMR_ERR_ARGUMENT, "flag buffer smaller than the leaf count");
                std::copy(r.flags.begin(), r.flags.end(), flags);
            }
        }
    );
}

int amr_tape_get_keys(const amr_tape* tape, uint64_t index, uint64_t* keys, uint64_t count, uint64_t* leaves)
{
    return guarded(
        [&]
        {
            if (!tape) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            if (index >= tape->topos.size()) throw amr_error(AMR_ERR_ARGUMENT, "topology index out of range");
            auto const& t = tape->topos[index];
            if (leaves) *leaves = t.n;
            if (keys)
            {
                if (count < t.n) throw amr_error(AMR_ERR_ARGUMENT, "key buffer smaller than the leaf count");
                AMR_HIP_CHECK(hipMemcpy(keys, t.keys, t.n * sizeof(uint64_t), hipMemcpyDeviceToHost));
            }
        }
    );
}

int amr_tape_forward(amr_tape* tape, const double* u0_device, const double velocity[3], double* uT_device)
{
    return guarded(
        [&]
        {
            if (!tape || !u0_device || !velocity || !uT_device) throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            tape->forward(u0_device, velocity, uT_device);
        }
    );
}

int amr_tape_adjoint(amr_tape* tape, const double* lambda_T_device, double* grad_u0_device, double grad_velocity[3])
{
    return guarded(
        [&]
        {
            if (!tape || !lambda_T_device || !grad_u0_device || !grad_velocity)
                throw amr_error(AMR_ERR_ARGUMENT, "null argument");
            tape->adjoint(lambda_T_device, grad_u0_device, grad_velocity);
        }
    );
}

} // extern "C"

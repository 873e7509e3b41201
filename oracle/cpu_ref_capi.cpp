// cpu_ref_capi.cpp -- extern "C" face of the CPU oracle (test infrastructure
// only; see cpu_ref.hpp).  Loaded with ctypes by tests/, smoke() and the
// cpu_baseline leg of bench.py.
#include "cpu_ref.hpp"

#include <memory>
#include <string>

namespace
{

thread_local std::string g_err;

struct iface
{
    virtual ~iface()                                                     = default;
    virtual int         rank() const                                      = 0;
    virtual int         interior_cells() const                            = 0;
    virtual int         padded_cells() const                              = 0;
    virtual std::size_t size() const                                      = 0;
    virtual void        keys(std::uint64_t* out) const                    = 0;
    virtual void        metadata(oracle::halo_meta* out) const            = 0;
    virtual bool        reconstruct(std::int8_t const* flags)             = 0;
    virtual void        last_sets(std::uint64_t* r, std::size_t* nr, std::uint64_t* c, std::size_t* nc) const = 0;
    virtual void        set_interior(double const* v)                     = 0;
    virtual void        get_interior(double* v) const                     = 0;
    virtual void        set_padded(double const* v)                       = 0;
    virtual void        get_padded(double* v) const                       = 0;
    virtual void        cell_centers(double* v) const                     = 0;
    virtual void        halo()                                            = 0;
    virtual void        params(double const* vel, double cfl, int threads) = 0;
    virtual double      dt() const                                        = 0;
    virtual void        step(double dt)                                   = 0;
    virtual void        step_kernel_only(double dt)                       = 0;
    virtual void        swap()                                            = 0;
    virtual double      advance(std::size_t steps, double rem, std::size_t* ex) = 0;
    virtual void        criterion(int kind, double r, double c, int mn, int mx, std::int8_t* f) const = 0;
};

template <int R, int H, int... Ns>
struct impl final : iface
{
    using tree_t = oracle::tree<R, H, Ns...>;
    tree_t t;

    impl(int depth, std::size_t cap, std::array<double, 3> L)
        : t(depth, cap, L)
    {
    }

    int         rank() const override { return R; }
    int         interior_cells() const override { return tree_t::NI; }
    int         padded_cells() const override { return tree_t::FLAT; }
    std::size_t size() const override { return t.size(); }
    void        keys(std::uint64_t* out) const override
    {
        for (std::size_t i = 0; i < t.size(); ++i) out[i] = t.key_at(i);
    }
    void metadata(oracle::halo_meta* out) const override { t.export_metadata(out); }
    bool reconstruct(std::int8_t const* flags) override { return t.reconstruct_tree(flags); }
    void last_sets(std::uint64_t* r, std::size_t* nr, std::uint64_t* c, std::size_t* nc) const override
    {
        auto const& a = t.last_refine_set();
        auto const& b = t.last_coarsen_set();
        if (r) std::copy(a.begin(), a.end(), r);
        if (c) std::copy(b.begin(), b.end(), c);
        if (nr) *nr = a.size();
        if (nc) *nc = b.size();
    }
    void set_interior(double const* v) override
    {
        for (std::size_t i = 0; i < t.size(); ++i) t.set_interior(i, v + i * tree_t::NI);
    }
    void get_interior(double* v) const override
    {
        for (std::size_t i = 0; i < t.size(); ++i) t.get_interior(i, v + i * tree_t::NI);
    }
    void set_padded(double const* v) override
    {
        for (std::size_t i = 0; i < t.size(); ++i) std::copy_n(v + i * tree_t::FLAT, tree_t::FLAT, t.patch(i));
    }
    void get_padded(double* v) const override
    {
        for (std::size_t i = 0; i < t.size(); ++i) std::copy_n(t.patch(i), tree_t::FLAT, v + i * tree_t::FLAT);
    }
    void cell_centers(double* v) const override
    {
        for (std::size_t i = 0; i < t.size(); ++i) t.cell_centers(i, v + i * tree_t::NI * R);
    }
    void   halo() override { t.halo_exchange_update(); }
    void   params(double const* vel, double cfl, int threads) override
    {
        t.set_velocity({ vel[0], vel[1], vel[2] });
        t.set_cfl(cfl);
        t.set_threads(threads);
    }
    double dt() const override { return t.compute_time_step(); }
    void   step(double dt) override { t.time_step(dt); }
    void   swap() override { t.swap_buffers(); }
    void   step_kernel_only(double dt) override
    {
        t.time_step_kernel(dt);
        t.swap_buffers();
    }
    double advance(std::size_t steps, double rem, std::size_t* ex) override { return t.advance_batch(steps, rem, ex); }
    void   criterion(int kind, double r, double c, int mn, int mx, std::int8_t* f) const override
    {
        if (kind == 0)
            t.criterion_max_value(r, c, mn, mx, f);
        else
            t.criterion_gradient(r, c, mn, mx, f);
    }
};

// Whether (rank, ext, h) is the instance <R, H, Ns...>.
template <int R, int H, int... Ns>
bool is(int rank, int const* ext, int h)
{
    constexpr int ns[] = { Ns... };
    if (rank != R || h != H) return false;
    for (int d = 0; d < R; ++d)
        if (ext[d] != ns[d]) return false;
    return true;
}

// Instances: the square / cubic shapes of the configurations and the
// reference's tests and benchmarks, small shapes for exhaustive tests, and
// non-square shapes (static_shape<N, M[, O]>) in both orientations.
#define ORACLE_SHAPES(X) \
    X(2, 1, 4, 4)        \
    X(2, 2, 4, 4)        \
    X(2, 1, 6, 6)        \
    X(2, 1, 8, 8)        \
    X(2, 1, 16, 16)      \
    X(2, 1, 64, 64)      \
    X(2, 2, 10, 10)      \
    X(2, 2, 50, 50)      \
    X(2, 1, 16, 8)       \
    X(2, 1, 8, 16)       \
    X(2, 2, 12, 8)       \
    X(3, 1, 4, 4, 4)     \
    X(3, 1, 8, 8, 8)     \
    X(3, 1, 4, 6, 8)     \
    X(3, 2, 6, 4, 4)

iface* make(int rank, int const* ext, int h, int depth, std::size_t cap, std::array<double, 3> L)
{
#define ORACLE_CASE(R_, H_, ...) \
    if (is<R_, H_, __VA_ARGS__>(rank, ext, h)) return new impl<R_, H_, __VA_ARGS__>(depth, cap, L);
    ORACLE_SHAPES(ORACLE_CASE)
#undef ORACLE_CASE
    throw std::invalid_argument("no oracle instance for this (rank, patch extents, halo)");
}

// The Euler system (several fields over one topology).
struct eiface
{
    virtual ~eiface()                                                      = default;
    virtual int         rank() const                                       = 0;
    virtual int         fields() const                                     = 0;
    virtual std::size_t size() const                                       = 0;
    virtual void        keys(std::uint64_t* out) const                     = 0;
    virtual void        metadata(oracle::halo_meta* out) const             = 0;
    virtual bool        reconstruct(std::int8_t const* flags)              = 0;
    virtual void        set_interior(double const* v)                      = 0;
    virtual void        get_interior(double* v) const                      = 0;
    virtual void        get_padded(double* v) const                        = 0;
    virtual void        cell_centers(double* v) const                      = 0;
    virtual void        halo()                                             = 0;
    virtual void        params(double gamma, double cfl, int threads)      = 0;
    virtual double      dt() const                                         = 0;
    virtual void        step(double dt)                                    = 0;
    virtual double      advance(std::size_t steps, double rem, std::size_t* ex) = 0;
    virtual void        criterion(int kind, double r, double c, int mn, int mx, std::int8_t* f) const = 0;
};

// Field-major host arrays: all leaves of field 0, then field 1, ...
template <int R, int H, int... Ns>
struct eimpl final : eiface
{
    using sys_t  = oracle::euler_system<R, H, Ns...>;
    using tree_t = typename sys_t::tree_t;
    sys_t e;

    eimpl(int depth, std::size_t cap, std::array<double, 3> L)
        : e(depth, cap, L)
    {
    }
    int         rank() const override { return R; }
    int         fields() const override { return sys_t::NV; }
    std::size_t size() const override { return e.size(); }
    void        keys(std::uint64_t* out) const override
    {
        for (std::size_t i = 0; i < e.size(); ++i) out[i] = e.field(0).key_at(i);
    }
    void metadata(oracle::halo_meta* out) const override { e.field(0).export_metadata(out); }
    bool reconstruct(std::int8_t const* flags) override { return e.reconstruct(flags); }
    void set_interior(double const* v) override
    {
        const std::size_t n = e.size();
        for (int f = 0; f < sys_t::NV; ++f)
            for (std::size_t i = 0; i < n; ++i) e.field(f).set_interior(i, v + (f * n + i) * tree_t::NI);
    }
    void get_interior(double* v) const override
    {
        const std::size_t n = e.size();
        for (int f = 0; f < sys_t::NV; ++f)
            for (std::size_t i = 0; i < n; ++i) e.field(f).get_interior(i, v + (f * n + i) * tree_t::NI);
    }
    void get_padded(double* v) const override
    {
        const std::size_t n = e.size();
        for (int f = 0; f < sys_t::NV; ++f)
            for (std::size_t i = 0; i < n; ++i)
                std::copy_n(e.field(f).patch(i), tree_t::FLAT, v + (f * n + i) * tree_t::FLAT);
    }
    void cell_centers(double* v) const override
    {
        for (std::size_t i = 0; i < e.size(); ++i) e.field(0).cell_centers(i, v + i * tree_t::NI * R);
    }
    void   halo() override { e.halo_exchange_update(); }
    void   params(double gamma, double cfl, int threads) override { e.set_params(gamma, cfl, threads); }
    double dt() const override { return e.compute_time_step(); }
    void   step(double dt) override { e.time_step(dt); }
    double advance(std::size_t steps, double rem, std::size_t* ex) override { return e.advance_batch(steps, rem, ex); }
    void   criterion(int kind, double r, double c, int mn, int mx, std::int8_t* f) const override
    {
        // the reference's Euler drivers flag on the density field
        // (benchmark/bench_fvm_solver_integration3D.b.cpp:86-107)
        if (kind == 0)
            e.field(0).criterion_max_value(r, c, mn, mx, f);
        else
            e.field(0).criterion_gradient(r, c, mn, mx, f);
    }
};

eiface* make_euler(int rank, int const* ext, int h, int depth, std::size_t cap, std::array<double, 3> L)
{
#define ORACLE_CASE(R_, H_, ...) \
    if (is<R_, H_, __VA_ARGS__>(rank, ext, h)) return new eimpl<R_, H_, __VA_ARGS__>(depth, cap, L);
    ORACLE_SHAPES(ORACLE_CASE)
#undef ORACLE_CASE
    throw std::invalid_argument("no Euler oracle instance for this (rank, patch extents, halo)");
}

template <typename Fn>
int guarded(Fn&& fn)
{
    try
    {
        fn();
        return 0;
    }
    catch (std::exception const& e)
    {
        g_err = e.what();
        return -1;
    }
}

} // namespace

extern "C"
{

const char* oracle_last_error(void) { return g_err.c_str(); }

void* oracle_create(int rank, int const* ext, int h, int depth, std::size_t capacity, double const* lengths)
{
    iface* p = nullptr;
    guarded([&] { p = make(rank, ext, h, depth, capacity, { lengths[0], lengths[1], lengths[2] }); });
    return p;
}

void oracle_destroy(void* o) { delete static_cast<iface*>(o); }

std::size_t oracle_size(void* o) { return static_cast<iface*>(o)->size(); }

int oracle_keys(void* o, std::uint64_t* out)
{
    return guarded([&] { static_cast<iface*>(o)->keys(out); });
}

int oracle_metadata(void* o, void* out)
{
    return guarded([&] { static_cast<iface*>(o)->metadata(static_cast<oracle::halo_meta*>(out)); });
}

int oracle_reconstruct(void* o, std::int8_t const* flags, int* changed)
{
    return guarded(
        [&]
        {
            const bool c = static_cast<iface*>(o)->reconstruct(flags);
            if (changed) *changed = c ? 1 : 0;
        }
    );
}

int oracle_last_sets(void* o, std::uint64_t* r, std::size_t* nr, std::uint64_t* c, std::size_t* nc)
{
    return guarded([&] { static_cast<iface*>(o)->last_sets(r, nr, c, nc); });
}

int oracle_set_interior(void* o, double const* v)
{
    return guarded([&] { static_cast<iface*>(o)->set_interior(v); });
}
int oracle_get_interior(void* o, double* v)
{
    return guarded([&] { static_cast<iface*>(o)->get_interior(v); });
}
int oracle_set_padded(void* o, double const* v)
{
    return guarded([&] { static_cast<iface*>(o)->set_padded(v); });
}
int oracle_get_padded(void* o, double* v)
{
    return guarded([&] { static_cast<iface*>(o)->get_padded(v); });
}
int oracle_cell_centers(void* o, double* v)
{
    return guarded([&] { static_cast<iface*>(o)->cell_centers(v); });
}
int oracle_halo_exchange(void* o)
{
    return guarded([&] { static_cast<iface*>(o)->halo(); });
}
int oracle_set_params(void* o, double const* vel, double cfl, int threads)
{
    return guarded([&] { static_cast<iface*>(o)->params(vel, cfl, threads); });
}
double oracle_compute_dt(void* o) { return static_cast<iface*>(o)->dt(); }
int    oracle_step(void* o, double dt)
{
    return guarded([&] { static_cast<iface*>(o)->step(dt); });
}
// ndtree::swap_buffers (ndtree.hpp:1558-1579): current <-> next.
int oracle_swap_buffers(void* o)
{
    return guarded([&] { static_cast<iface*>(o)->swap(); });
}
int oracle_step_kernel_only(void* o, double dt)
{
    return guarded([&] { static_cast<iface*>(o)->step_kernel_only(dt); });
}
int oracle_advance_batch(void* o, std::size_t steps, double remaining, double* dt_sum, std::size_t* executed)
{
    return guarded([&] { *dt_sum = static_cast<iface*>(o)->advance(steps, remaining, executed); });
}
int oracle_criterion(void* o, int kind, double r, double c, int mn, int mx, std::int8_t* flags)
{
    return guarded([&] { static_cast<iface*>(o)->criterion(kind, r, c, mn, mx, flags); });
}

// ---- Euler system -------------------------------------------------------

void* oracle_euler_create(int rank, int const* ext, int h, int depth, std::size_t capacity, double const* lengths)
{
    eiface* p = nullptr;
    guarded([&] { p = make_euler(rank, ext, h, depth, capacity, { lengths[0], lengths[1], lengths[2] }); });
    return p;
}
void        oracle_euler_destroy(void* o) { delete static_cast<eiface*>(o); }
std::size_t oracle_euler_size(void* o) { return static_cast<eiface*>(o)->size(); }
int         oracle_euler_fields(void* o) { return static_cast<eiface*>(o)->fields(); }
int         oracle_euler_keys(void* o, std::uint64_t* out)
{
    return guarded([&] { static_cast<eiface*>(o)->keys(out); });
}
int oracle_euler_metadata(void* o, void* out)
{
    return guarded([&] { static_cast<eiface*>(o)->metadata(static_cast<oracle::halo_meta*>(out)); });
}
int oracle_euler_reconstruct(void* o, std::int8_t const* flags, int* changed)
{
    return guarded(
        [&]
        {
            const bool c = static_cast<eiface*>(o)->reconstruct(flags);
            if (changed) *changed = c ? 1 : 0;
        }
    );
}
int oracle_euler_set_interior(void* o, double const* v)
{
    return guarded([&] { static_cast<eiface*>(o)->set_interior(v); });
}
int oracle_euler_get_interior(void* o, double* v)
{
    return guarded([&] { static_cast<eiface*>(o)->get_interior(v); });
}
int oracle_euler_get_padded(void* o, double* v)
{
    return guarded([&] { static_cast<eiface*>(o)->get_padded(v); });
}
int oracle_euler_cell_centers(void* o, double* v)
{
    return guarded([&] { static_cast<eiface*>(o)->cell_centers(v); });
}
int oracle_euler_halo_exchange(void* o)
{
    return guarded([&] { static_cast<eiface*>(o)->halo(); });
}
int oracle_euler_set_params(void* o, double gamma, double cfl, int threads)
{
    return guarded([&] { static_cast<eiface*>(o)->params(gamma, cfl, threads); });
}
double oracle_euler_compute_dt(void* o) { return static_cast<eiface*>(o)->dt(); }
int    oracle_euler_step(void* o, double dt)
{
    return guarded([&] { static_cast<eiface*>(o)->step(dt); });
}
int oracle_euler_advance_batch(void* o, std::size_t steps, double remaining, double* dt_sum, std::size_t* executed)
{
    return guarded([&] { *dt_sum = static_cast<eiface*>(o)->advance(steps, remaining, executed); });
}
int oracle_euler_criterion(void* o, int kind, double r, double c, int mn, int mx, std::int8_t* flags)
{
    return guarded([&] { static_cast<eiface*>(o)->criterion(kind, r, c, mn, mx, flags); });
}

// Pointwise Euler functions (for the closed-form checks).
int oracle_euler_prim_to_cons(int rank, double const* prim, double* cons, std::size_t cells, double gamma)
{
    return guarded(
        [&]
        {
            const int nv = rank + 2;
            for (std::size_t c = 0; c < cells; ++c)
            {
                if (rank == 2)
                    oracle::euler_physics<2>::prim_to_cons(prim + c * nv, cons + c * nv, gamma);
                else if (rank == 3)
                    oracle::euler_physics<3>::prim_to_cons(prim + c * nv, cons + c * nv, gamma);
                else
                    throw std::invalid_argument("rank must be 2 or 3");
            }
        }
    );
}
int oracle_euler_flux(int rank, double const* UL, double const* UR, double* flux, int dir, double gamma)
{
    return guarded(
        [&]
        {
            if (rank == 2)
                oracle::euler_physics<2>::rusanov(UL, UR, flux, dir, gamma);
            else if (rank == 3)
                oracle::euler_physics<3>::rusanov(UL, UR, flux, dir, gamma);
            else
                throw std::invalid_argument("rank must be 2 or 3");
        }
    );
}
double oracle_advection_flux(double ul, double ur, double v) { return oracle::advection_flux(ul, ur, v); }

double oracle_euler_max_speed(int rank, double const* U, int dir, double gamma)
{
    return rank == 2 ? oracle::euler_physics<2>::max_speed(U, dir, gamma) : oracle::euler_physics<3>::max_speed(U, dir, gamma);
}

} // extern "C"

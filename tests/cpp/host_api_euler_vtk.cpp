// The reference's 3D Euler benchmark loop (benchmark/bench_fvm_solver_integration3D.b.cpp:
// 8^3 patches, morton_id<5, 3>, domain 1000^3, gamma 1.4, cfl 0.3, two
// refine-all, acoustic pulse, regrid every 5 steps by maximum density) run
// through the amr::hip C++ host API for a few cycles, then a VTK file written
// with vtk_print.  Checks: the solver advances (positive dt, 5 steps per
// cycle), mass and energy are conserved while the tree is uniform, the
// criterion refines around the pulse, the per-field padded patch view works,
// and the VTK file has the expected size.  Prints PASS/FAIL per check; the
// exit status counts failures.
#include "amr_hip/ndtree.hpp"

#include <cmath>
#include <cstdio>
#include <filesystem>

namespace
{
int failures = 0;

void expect(bool ok, char const* what)
{
    std::printf("  %s: %s\n", what, ok ? "ok" : "FAIL");
    if (!ok) ++failures;
}

template <typename Tree>
std::array<double, 5> totals(Tree const& tree)
{
    const auto            u   = tree.get_interior();
    const std::size_t     per = tree.size() * Tree::s_interior;
    std::array<double, 5> t{};
    for (unsigned f = 0; f < 5; ++f)
        for (std::size_t i = 0; i < per; ++i) t[f] += u[f * per + i];
    return t;
}
} // namespace

int main()
{
    try
    {
        using tree_t = amr::hip::ndtree<3, 5, 8, 1>;
        constexpr double L = 1000.0;
        amr::hip::euler_solver<tree_t> solver(50000, 1.4, 0.3, { L, L, L });
        auto& tree = solver.get_tree();
        for (int i = 0; i < 2; ++i) tree.reconstruct_tree([](auto const&) { return amr::hip::refine_status_t::Refine; });
        expect(tree.size() == 64, "two refine-all give 64 leaves");
        expect(tree.fields() == 5, "five fields");
        solver.initialize(
            [&](std::array<double, 3> const& c)
            {
                double r2 = 0.0;
                for (int d = 0; d < 3; ++d) r2 += (c[d] - 0.5 * L) * (c[d] - 0.5 * L);
                const double pert = 10.0 * std::exp(-r2 / (0.01 * L * L));
                return std::array<double, 5>{ 0.5 + pert * 0.2, 0.0, 0.0, 0.0, 1.0 + pert };
            }
        );
        const auto t0 = totals(tree);
        std::size_t executed = 0;
        solver.advance_batch_async(5, 200.0);
        const double dt5 = solver.finish_advance_batch(&executed);
        expect(executed == 5 && dt5 > 0.0, "five steps with positive dt");
        const auto t1 = totals(tree);
        expect(std::abs(t1[0] - t0[0]) <= 1e-12 * t0[0], "mass conserved on the uniform tree");
        expect(std::abs(t1[4] - t0[4]) <= 1e-12 * t0[4], "energy conserved on the uniform tree");
        const auto crit = amr::hip::device_criterion::max_value(0.53, 0.49, 1, 5);
        double     t    = dt5;
        for (int cycle = 0; cycle < 3; ++cycle)
        {
            tree.reconstruct_tree(crit);
            solver.advance_batch_async(5, 200.0 - t);
            t += solver.finish_advance_batch(&executed);
        }
        expect(tree.size() > 64, "density criterion refined around the pulse");
        const auto        rho = tree.get_patch(std::size_t{ 0 }, 0);
        const auto        e   = tree.get_patch(std::size_t{ 0 }, 4);
        constexpr auto    P   = tree_t::s_padded;
        const std::size_t mid = (P[0] / 2 * P[1] + P[1] / 2) * P[2] + P[2] / 2; // an interior cell
        expect(rho.size() == tree_t::s_flat && rho[mid] > 0.0 && e[mid] > rho[mid], "per-field padded patches");
        amr::hip::vtk_print vtk("host_api_euler");
        vtk.print(tree, ".vtk");
        const auto bytes = std::filesystem::file_size("vtk_output/host_api_euler.vtk");
        const auto cells = tree.size() * tree_t::s_interior;
        // points (8 x 3 doubles), connectivity (9 ints), type, index, halo
        // flag (ints) and five double fields per cell, plus the text lines
        const auto payload = cells * (8 * 3 * 8 + 9 * 4 + 4 + 4 + 4 + 5 * 8);
        expect(bytes > payload && bytes < payload + 1024, "VTK file size");
        std::printf("leaves %zu, t = %.3f\n", tree.size(), t);
    }
    catch (std::exception const& ex)
    {
        std::printf("exception: %s\n", ex.what());
        return 2;
    }
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "PASSED", failures);
    return failures ? 1 : 0;
}

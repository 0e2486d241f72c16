// Golden-vector harness: drives the reference solver's own public functions
// in the order of Solver::initialise (src/solver/Solver.cpp:242-302),
// Solver::preparePressureSystem (:587-593) and Solver::advanceOneStep
// (:609-680), and dumps the state after every stage.  Built by oracle/ref.mk
// against objects compiled from the reference's sources; nothing of the
// reference is copied here.  Test infrastructure only.
//
//   ref_dump <case_base> <out.dump> <nsteps> [cg_trace_iters] [post]
//
// With "post", the reference's own output routines also run, in the order of
// Solver::run (src/solver/Solver.cpp:150, 177-183, 211-221): the residual
// CSV, VTU and PVD files go to output/<case>/ under the working directory,
// and the time and solution arrays they were written from are dumped, with
// the wall-clock seconds spent in Post::writeSolution ("post_seconds", its
// calls counted in "post_solutions").
//
// Output: a flat binary file of records
//   int32 name_len, name bytes, char dtype ('d' float64 | 'i' int32),
//   int64 count, count elements.

#include "cbs/assembly/EnergyAssembly.hpp"
#include "cbs/assembly/MomentumAssembly.hpp"
#include "cbs/assembly/PressureAssembly.hpp"
#include "cbs/boundary/Boundary.hpp"
#include "cbs/core/CBSStateSI.hpp"
#include "cbs/io/MeshIO.hpp"
#include "cbs/io/Post.hpp"
#include "cbs/linalg/ConjugateGradient.hpp"
#include "cbs/linalg/MatrixVectorCalc.hpp"
#include "cbs/parallel/Coloring.hpp"
#include "cbs/preprocess/Preprocess.hpp"
#include "cbs/solver/Convergence.hpp"
#include "cbs/solver/Steps.hpp"
#include "cbs/timestep/TimeStep.hpp"

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <string>
#include <vector>

using namespace cbs;

namespace
{
    FILE* g_out = nullptr;

    void put(const std::string& name, char dtype, const void* data, std::int64_t count)
    {
        const std::int32_t len = static_cast<std::int32_t>(name.size());
        std::fwrite(&len, sizeof(len), 1, g_out);
        std::fwrite(name.data(), 1, name.size(), g_out);
        std::fwrite(&dtype, 1, 1, g_out);
        std::fwrite(&count, sizeof(count), 1, g_out);
        const std::size_t esz = (dtype == 'd') ? 8 : 4;
        if (count > 0)
        {
            std::fwrite(data, esz, static_cast<std::size_t>(count), g_out);
        }
    }

    // Arrays are 1-based with element 0 unused: dump elements 1..n.
    template <typename A>
    void put_real(const std::string& name, const A& a, std::int64_t n = -1)
    {
        const auto& v = a.raw();
        const std::int64_t cnt = (n < 0) ? static_cast<std::int64_t>(v.size()) - 1 : n;
        put(name, 'd', v.data() + 1, cnt);
    }

    template <typename A>
    void put_int(const std::string& name, const A& a, std::int64_t n = -1)
    {
        const auto& v = a.raw();
        const std::int64_t cnt = (n < 0) ? static_cast<std::int64_t>(v.size()) - 1 : n;
        put(name, 'i', v.data() + 1, cnt);
    }

    void put_scalar(const std::string& name, double v) { put(name, 'd', &v, 1); }
    void put_iscalar(const std::string& name, std::int32_t v) { put(name, 'i', &v, 1); }

    void update_velocity_magnitude(CBSStateSI& s)
    {
        // Same expression as Solver::updateVelocityMagnitude (Solver.cpp:689-699).
        for (Int ip = 1; ip <= s.cfg.npoin; ++ip)
        {
            const Real u = s.unkno(1, ip);
            const Real v = s.unkno(2, ip);
            const Real w = s.unkno(3, ip);
            s.velocity(ip) = std::sqrt(u * u + v * v + w * w + 1.0e-16);
        }
    }

    void put_cg(const std::string& p, const ConjugateGradient::Result& r)
    {
        put_iscalar(p + "cg_iterations", r.iterations);
        put_scalar(p + "cg_initial_l2", r.initial_l2);
        put_scalar(p + "cg_final_l2", r.final_l2);
        put_scalar(p + "cg_relative_l2", r.final_relative_l2);
        put_scalar(p + "cg_max_abs", r.final_max_abs);
        put_iscalar(p + "cg_converged", r.converged ? 1 : 0);
    }
}

int main(int argc, char** argv)
{
    if (argc < 4)
    {
        std::fprintf(stderr, "usage: ref_dump <case_base> <out.dump> <nsteps> [cg_trace_iters] [post]\n");
        return 2;
    }
    const std::string case_name = argv[1];
    const int nsteps = std::atoi(argv[3]);
    const int trace_iters = (argc > 4) ? std::atoi(argv[4]) : 0;
    const bool post = (argc > 5) && std::string(argv[5]) == "post";
    const std::string tag = std::filesystem::path(case_name).filename().string();
    g_out = std::fopen(argv[2], "wb");
    if (!g_out)
    {
        std::perror("fopen");
        return 2;
    }

    try
    {
        CBSStateSI s;
        MeshIO::readAll(case_name, s);
        Preprocess::validateBoundaryFlags(s);
        Preprocess::shapeFunctionDerivatives(s);
        Preprocess::assignBoundaryFaceNumbers(s);
        Preprocess::getNormals(s);
        Preprocess::massMatrix(s);
        Preprocess::classifyFaceEdges(s);
        Preprocess::elementSize(s);
        Preprocess::wallDetermination(s);
        Preprocess::computeMassFlowInletVelocity(s);
        Preprocess::initialiseVelocityMagnitude(s);
        Preprocess::detectPressureBoundaryNodes(s);
        Coloring::build(s);
        Boundary::applyTemperature(s);
        Boundary::applyVelocity(s);
        Boundary::applyPressure(s);
        update_velocity_magnitude(s);

        put_iscalar("npoin", s.cfg.npoin);
        put_iscalar("nelem", s.cfg.nelem);
        put_iscalar("nboun", s.cfg.nboun);
        put_scalar("ani", s.cfg.ani);
        put_scalar("inlet_u_from_massflow", s.cfg.inlet_u_from_massflow);
        put_iscalar("temp_calc", s.cfg.temp_calc);
        put_int("intma", s.intma);
        put_real("coord", s.coord);
        put_int("mat_elem", s.mat_elem);
        put_real("rho_e", s.rho_e);
        put_real("mu_e", s.mu_e);
        put_real("k_e", s.k_e);
        put_real("rho_cp_e", s.rho_cp_e);
        put_real("alpha_e", s.alpha_e);
        put_real("Qvol_e", s.Qvol_e);
        put_real("dNkdx", s.dNkdx);
        put_real("detJ", s.detJ);
        put_real("annxf", s.annxf);
        put_int("fedge", s.fedge);
        put_int("iside", s.iside);
        put_real("face_norm", s.face_norm);
        put_real("elcoe_e", s.elcoe_e);
        put_real("alen_e", s.alen_e);
        put_real("Mdiag_real", s.Mdiag_real);
        put_iscalar("npoin_wall", s.cfg.npoin_wall);
        put_int("wall_node_list", s.wall_node_list, s.cfg.npoin_wall);
        put_iscalar("bc_fixed", s.cfg.bc_fixed);
        put_int("bc_list", s.bc_list, s.cfg.bc_fixed);
        put_real("bc_values", s.bc_values, s.cfg.bc_fixed);
        put_iscalar("ncolor", s.ncolor);
        put("color_ptr", 'i', s.color_ptr.data(), static_cast<std::int64_t>(s.color_ptr.size()));
        put("color_elem", 'i', s.color_elem.data(), static_cast<std::int64_t>(s.color_elem.size()));
        put_real("init/unkno", s.unkno);
        put_real("init/pres", s.pres);
        put_real("init/temperature", s.temperature);
        put_real("init/velocity", s.velocity);

        double post_seconds = 0.0;
        int post_solutions = 0;
        auto timed_solution = [&](int iitime) {
            const auto t0 = std::chrono::steady_clock::now();
            Post::writeSolution(s, tag, iitime);
            post_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            ++post_solutions;
        };
        if (post)
        {
            put_scalar("init/rtime", s.cfg.rtime);
            // Post::initialiseRunOutputs writes the iteration-0 solution
            // itself when VTU output is enabled
            const auto t0 = std::chrono::steady_clock::now();
            Post::initialiseRunOutputs(s, tag);
            if (s.cfg.vtu_output_enabled > 0)
            {
                post_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                ++post_solutions;
            }
        }

        PressureAssembly::buildElementPressureTerms(s);
        PressureAssembly::buildGlobalPressureTerms(s);
        put_real("pdiagE", s.pdiagE);
        put_real("gstifE", s.gstifE);

        for (int it = 1; it <= nsteps; ++it)
        {
            const std::string p = "s" + std::to_string(it) + "/";
            s.unkn1 = s.unkno;
            s.pres1 = s.pres;
            s.temperature1 = s.temperature;

            TimeStep::computeTimeStep(s, it);
            if (s.cfg.transient_on > 0)
            {
                s.cfg.rtime += s.cfg.dtreal;
            }
            put_scalar(p + "dtreal", s.cfg.dtreal);
            put_real(p + "delte", s.delte);
            put_real(p + "deltp", s.deltp);
            put_real(p + "deltp1", s.deltp1);
            put_real(p + "deltp2", s.deltp2);
            put_real(p + "beta1", s.beta1);
            put_real(p + "vvis", s.vvis);
            put_real(p + "velcp", s.velcp);
            put_real(p + "unkno_unit_vec", s.unkno_unit_vec);
            put_real(p + "alen", s.alen);

            TimeStep::updateLhsDiagonal(s);
            put_real(p + "elcoe2", s.elcoe2);
            put_real(p + "elcoe2p", s.elcoe2p);
            put_real(p + "lhs_pdiag", s.pdiag);
            put_real(p + "gstif", s.gstif);
            put_real(p + "lhs_pres", s.pres);
            put_real(p + "lhs_rhs1", s.rhs1);

            Steps::step1(s);
            put_real(p + "step1_rhs", s.rhs);
            put_real(p + "step1_unkno", s.unkno);

            PressureAssembly::assembleStep2Rhs(s);
            put_real(p + "step2_rhs1", s.rhs1);
            if (it == 1)
            {
                // Per-iteration CG residual history: restart from the same
                // state with cg_max_iter = k.
                if (trace_iters > 0)
                {
                    std::vector<double> l2(trace_iters), mx(trace_iters);
                    for (int k = 1; k <= trace_iters; ++k)
                    {
                        CBSStateSI t = s;
                        t.cfg.cg_max_iter = k;
                        const auto r = ConjugateGradient::solvePressure(t);
                        l2[k - 1] = r.final_l2;
                        mx[k - 1] = r.final_max_abs;
                    }
                    put(p + "cg_trace_l2", 'd', l2.data(), trace_iters);
                    put(p + "cg_trace_max", 'd', mx.data(), trace_iters);
                }
                // One matvec of the pre-solve pressure with the step's operator.
                Array1D<Real> y;
                y.resize(s.cfg.npoin);
                MatrixVectorCalc::pressureMultiply(s, s.pres, y);
                put_real(p + "apply_x", s.pres);
                put_real(p + "apply_y", y);
            }
            Steps::step2(s);
            put_iscalar(p + "cg_iterations", s.last_cg_iterations);
            put_scalar(p + "cg_initial_l2", s.last_cg_initial_l2);
            put_scalar(p + "cg_final_l2", s.last_cg_final_l2);
            put_scalar(p + "cg_relative_l2", s.last_cg_relative_l2);
            put_scalar(p + "cg_max_abs", s.last_cg_max_abs);
            put_real(p + "step2_pres", s.pres);
            put_real(p + "step2_rhs1_after", s.rhs1);
            put_real(p + "step2_pdiag_after", s.pdiag);

            Steps::step3(s);
            put_real(p + "step3_rhs", s.rhs);
            put_real(p + "step3_unkno", s.unkno);

            Steps::step4(s);
            put_real(p + "step4_rhs1", s.rhs1);
            put_real(p + "step4_temperature", s.temperature);

            update_velocity_magnitude(s);
            put_real(p + "velocity", s.velocity);
            Convergence::evaluate(s);
            put(p + "hb", 'd', s.hb.data(), 15);
            if (post)
            {
                put_scalar(p + "rtime", s.cfg.rtime);
                put_real(p + "post_pres", s.pres);
                put_real(p + "post_unkno", s.unkno);
                put_real(p + "post_temperature", s.temperature);
                put_real(p + "post_velocity", s.velocity);
                Post::writeResidualRow(s, tag, it);
                if (Post::shouldWriteVTU(s, it))
                {
                    timed_solution(it);
                }
            }
        }
        if (post && s.cfg.vtu_output_enabled > 0 &&
            !(!s.output_iterations.empty() && s.output_iterations.back() == nsteps))
        {
            timed_solution(nsteps);
        }
        if (post)
        {
            put_scalar("post_seconds", post_seconds);
            put_iscalar("post_solutions", post_solutions);
        }
    }
    catch (const std::exception& e)
    {
        std::fprintf(stderr, "ref_dump: %s\n", e.what());
        std::fclose(g_out);
        return 1;
    }
    std::fclose(g_out);
    return 0;
}

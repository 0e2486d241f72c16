// Host-side case set-up for the CBS hot path: the CBSStateSI arrays a run
// needs before its first step (see include/cbs_case.h).
//
// C++20.  Each routine restates the reference routine cited beside it with
// the same floating-point operation order, so the geometry handed to the GPU
// is bit-identical to what the reference's own initialisation produces.  The
// element-independent loops run under OpenMP; loops that accumulate into
// nodes stay in element order.

#include "cbs_case.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include <sys/stat.h>

namespace cbs_host
{
    using Real = double;
    using Int = std::int32_t;

    // Local face f is opposite local node f (CBSStateSI.hpp:342-365).
    constexpr int kIppn1[4][3] = {{3, 2, 4}, {3, 4, 1}, {1, 4, 2}, {1, 2, 3}};

    bool supported_bc(Int bc)
    {
        switch (bc)
        {
        case 500: case 501: case 502: case 503: case 504: case 506: case 507: case 508:
        case 510: case 511: case 520: case 530: case 532: case 901: case 902:
            return true;
        default:
            return false;
        }
    }

    // Preprocess.cpp:80-104
    bool is_no_slip_wall_bc(Int bc)
    {
        return bc == 501 || bc == 502 || bc == 530 || bc == 532 || bc == 901;
    }

    bool is_prescribed_velocity_face_bc(Int bc)
    {
        return bc == 500 || bc == 501 || bc == 502 || bc == 503 || bc == 507 || bc == 508 ||
               bc == 510 || bc == 511 || bc == 530 || bc == 532 || bc == 901;
    }

    struct Case
    {
        cbs_config cfg{};
        Int npoin = 0, nelem = 0, nboun = 0;
        Int ncolor = 0, npoin_wall = 0, bc_fixed = 0;
        std::vector<Int> intma, iside, mat_elem, fedge, wall_node_list, bc_list;
        std::vector<Int> color_ptr, color_elem;
        std::vector<Real> coord, unkno, pres, temperature;
        std::vector<Real> rho_e, cp_e, k_e, mu_e, rho_cp_e, alpha_e, Qvol_e;
        std::vector<Real> dNkdx, detJ, annxf, face_norm, elcoe_e, alen_e, Mdiag_real;
        std::vector<Real> wall_node_norm, bc_values;

        Int& IN(Int a, Int ie) { return intma[(std::size_t(ie) - 1) * 4 + a - 1]; }
        Int& IS(Int k, Int ib) { return iside[(std::size_t(ib) - 1) * 6 + k - 1]; }
        Real& X(Int d, Int ip) { return coord[(std::size_t(ip) - 1) * 3 + d - 1]; }
        Real& DN(Int ie, Int d, Int a) { return dNkdx[(std::size_t(ie) - 1) * 12 + (d - 1) * 4 + a - 1]; }
        Real& ANN(Int c, Int is, Int ie) { return annxf[((std::size_t(ie) - 1) * 4 + is - 1) * 4 + c - 1]; }
        Real& FN(Int k, Int ib) { return face_norm[(std::size_t(ib) - 1) * 4 + k - 1]; }

        void allocate()
        {
            const std::size_t E = std::size_t(nelem), N = std::size_t(npoin), B = std::size_t(nboun);
            intma.assign(4 * E, 0);
            iside.assign(6 * B, 0);
            mat_elem.assign(E, 0);
            fedge.assign(4 * E, 0);
            coord.assign(3 * N, 0.0);
            unkno.assign(3 * N, 0.0);
            pres.assign(N, 0.0);
            temperature.assign(N, 0.0);
            rho_e.assign(E, 1.0);
            cp_e.assign(E, 1.0);
            k_e.assign(E, 1.0);
            mu_e.assign(E, 0.0);
            rho_cp_e.assign(E, 1.0);
            alpha_e.assign(E, 1.0);
            Qvol_e.assign(E, 0.0);
            dNkdx.assign(12 * E, 0.0);
            detJ.assign(E, 0.0);
            annxf.assign(16 * E, 0.0);
            face_norm.assign(4 * B, 0.0);
            elcoe_e.assign(4 * E, 0.0);
            alen_e.assign(E, 0.0);
            Mdiag_real.assign(N, 0.0);
            wall_node_list.assign(N, 0);
            wall_node_norm.assign(3 * N, 0.0);
            bc_list.assign(N, 0);
            bc_values.assign(N, 0.0);
        }

        void touch_masks(std::vector<char>& tf, std::vector<char>& ts) const
        {
            tf.assign(std::size_t(npoin) + 1, 0);
            ts.assign(std::size_t(npoin) + 1, 0);
            for (Int ie = 1; ie <= nelem; ++ie)
            {
                const bool fluid = mat_elem[ie - 1] == 0;
                for (Int a = 0; a < 4; ++a)
                {
                    const Int ip = intma[(std::size_t(ie) - 1) * 4 + a];
                    (fluid ? tf : ts)[ip] = 1;
                }
            }
        }

        // MeshIO::initialiseFields and the material branches of readAll
        // (MeshIO.cpp:109-123, 471-570, 1627-1701).
        void initialise_fields(const Int* mats_elem, Int nmat, const cbs_material* mats)
        {
            if (cfg.material_properties_enabled > 0)
            {
                if (!mats_elem || !mats || nmat < 1)
                {
                    throw std::runtime_error("cbs_case - material files required by the .par controls");
                }
                std::unordered_map<Int, cbs_material> table;
                for (Int i = 0; i < nmat; ++i)
                {
                    const cbs_material& m = mats[i];
                    if (m.rho <= 0.0 || m.cp <= 0.0 || m.k <= 0.0)
                    {
                        throw std::runtime_error("MeshIO::readMaterialPropertyFile - material " +
                                                 std::to_string(m.id) + " has non-positive rho/cp/k");
                    }
                    if (m.mu < 0.0)
                    {
                        throw std::runtime_error("MeshIO::readMaterialPropertyFile - material " +
                                                 std::to_string(m.id) + " has negative viscosity");
                    }
                    if (!table.emplace(m.id, m).second)
                    {
                        throw std::runtime_error("MeshIO::readMaterialPropertyFile - duplicate material ID");
                    }
                }
                for (Int ie = 1; ie <= nelem; ++ie)
                {
                    mat_elem[ie - 1] = mats_elem[ie - 1];
                    const auto it = table.find(mat_elem[ie - 1]);
                    if (it == table.end())
                    {
                        throw std::runtime_error("MeshIO::readMaterialPropertyFile - element " +
                                                 std::to_string(ie) + " uses material ID " +
                                                 std::to_string(mat_elem[ie - 1]) + " missing from .matprop");
                    }
                    const cbs_material& m = it->second;
                    rho_e[ie - 1] = m.rho;
                    cp_e[ie - 1] = m.cp;
                    k_e[ie - 1] = m.k;
                    mu_e[ie - 1] = m.mu;
                    Qvol_e[ie - 1] = m.qvol;
                    rho_cp_e[ie - 1] = rho_e[ie - 1] * cp_e[ie - 1];
                    alpha_e[ie - 1] = k_e[ie - 1] / rho_cp_e[ie - 1];
                }
            }
            // set_freestream_initial_condition (MeshIO.cpp:109-123)
            for (Int ip = 1; ip <= npoin; ++ip)
            {
                for (Int d = 0; d < 3; ++d)
                {
                    unkno[(std::size_t(ip) - 1) * 3 + d] = cfg.cinf[d];
                }
                pres[ip - 1] = cfg.cinf[3];
                temperature[ip - 1] = cfg.cinf[4];
            }
            if (cfg.material_properties_enabled > 0)
            {
                // enforce_initial_material_domain_fields (MeshIO.cpp:471-570)
                std::vector<char> tf, ts;
                touch_masks(tf, ts);
                for (Int ip = 1; ip <= npoin; ++ip)
                {
                    if (!tf[ip] && !ts[ip])
                    {
                        throw std::runtime_error(
                            "MeshIO::initialiseFields - orphan node not touched by any material element: " +
                            std::to_string(ip));
                    }
                    if (ts[ip])
                    {
                        for (Int d = 0; d < 3; ++d) unkno[(std::size_t(ip) - 1) * 3 + d] = 0.0;
                    }
                    if (ts[ip] && !tf[ip]) pres[ip - 1] = 0.0;
                }
            }
            if (cfg.convection_type == 1)
            {
                cfg.ani = cfg.pr;
                cfg.temp_calc = 1;
            }
            else
            {
                cfg.ani = 1.0 / cfg.re;
                if (cfg.ri > 0.0) cfg.temp_calc = 1;
            }
            if (cfg.material_properties_enabled < 1)
            {
                for (Int ie = 1; ie <= nelem; ++ie)
                {
                    mat_elem[ie - 1] = 0;
                    rho_e[ie - 1] = 1.0;
                    cp_e[ie - 1] = 1.0;
                    rho_cp_e[ie - 1] = 1.0;
                    mu_e[ie - 1] = cfg.ani;
                    k_e[ie - 1] = 1.0;
                    alpha_e[ie - 1] = 1.0;
                    Qvol_e[ie - 1] = 0.0;
                }
            }
        }

        // Preprocess::validateBoundaryFlags (Preprocess.cpp:288-317) only
        // warns; Boundary's validate_boundary_flags throws on the first
        // boundary application (Boundary.cpp:70-92), so reject here.
        void validate_boundary_flags()
        {
            for (Int ib = 1; ib <= nboun; ++ib)
            {
                const Int bc = IS(6, ib);
                if (!supported_bc(bc))
                {
                    throw std::runtime_error("Boundary - unsupported BC_ID " + std::to_string(bc) +
                                             " at boundary face " + std::to_string(ib));
                }
            }
        }

        // Preprocess::shapeFunctionDerivatives (Preprocess.cpp:335-445)
        void shape_function_derivatives()
        {
            Int bad = 0;
#pragma omp parallel for schedule(static) reduction(max : bad)
            for (Int ie = 1; ie <= nelem; ++ie)
            {
                const Int* n = &intma[(std::size_t(ie) - 1) * 4];
                const Real* p1 = &coord[(std::size_t(n[0]) - 1) * 3];
                const Real* p2 = &coord[(std::size_t(n[1]) - 1) * 3];
                const Real* p3 = &coord[(std::size_t(n[2]) - 1) * 3];
                const Real* p4 = &coord[(std::size_t(n[3]) - 1) * 3];
                const Real j11 = p2[0] - p1[0], j21 = p2[1] - p1[1], j31 = p2[2] - p1[2];
                const Real j12 = p3[0] - p1[0], j22 = p3[1] - p1[1], j32 = p3[2] - p1[2];
                const Real j13 = p4[0] - p1[0], j23 = p4[1] - p1[1], j33 = p4[2] - p1[2];
                const Real det = j11 * (j22 * j33 - j23 * j32) - j12 * (j21 * j33 - j23 * j31) +
                                 j13 * (j21 * j32 - j22 * j31);
                if (det <= 0.0 || !std::isfinite(det))
                {
                    bad = std::max(bad, ie);
                    continue;
                }
                const Real inv = 1.0 / det;
                Real* g = &dNkdx[(std::size_t(ie) - 1) * 12];
                // g[(dim-1)*4 + (a-1)]
                const Real gx2 = (j22 * j33 - j23 * j32) * inv;
                const Real gy2 = (j13 * j32 - j12 * j33) * inv;
                const Real gz2 = (j12 * j23 - j13 * j22) * inv;
                const Real gx3 = (j23 * j31 - j21 * j33) * inv;
                const Real gy3 = (j11 * j33 - j13 * j31) * inv;
                const Real gz3 = (j13 * j21 - j11 * j23) * inv;
                const Real gx4 = (j21 * j32 - j22 * j31) * inv;
                const Real gy4 = (j12 * j31 - j11 * j32) * inv;
                const Real gz4 = (j11 * j22 - j12 * j21) * inv;
                g[0] = -(gx2 + gx3 + gx4);
                g[4] = -(gy2 + gy3 + gy4);
                g[8] = -(gz2 + gz3 + gz4);
                g[1] = gx2; g[5] = gy2; g[9] = gz2;
                g[2] = gx3; g[6] = gy3; g[10] = gz3;
                g[3] = gx4; g[7] = gy4; g[11] = gz4;
                detJ[ie - 1] = det;
            }
            if (bad)
            {
                // report the first failing element, as the serial loop would
                for (Int ie = 1; ie <= nelem; ++ie)
                {
                    if (detJ[ie - 1] <= 0.0 || !std::isfinite(detJ[ie - 1]))
                    {
                        throw std::runtime_error(
                            "Preprocess::shapeFunctionDerivatives - non-positive tetrahedron volume at element " +
                            std::to_string(ie) +
                            ". Tetrahedral connectivity must be consistently positively oriented.");
                    }
                }
            }
        }

        // Preprocess::assignBoundaryFaceNumbers (Preprocess.cpp:458-514)
        void assign_boundary_face_numbers()
        {
            for (Int ib = 1; ib <= nboun; ++ib)
            {
                const Int parent = IS(5, ib);
                if (parent < 1 || parent > nelem)
                {
                    throw std::runtime_error(
                        "Preprocess::assignBoundaryFaceNumbers - invalid parent tetrahedron at boundary face " +
                        std::to_string(ib));
                }
                std::array<Int, 3> bf = {IS(1, ib), IS(2, ib), IS(3, ib)};
                std::sort(bf.begin(), bf.end());
                Int matched = 0;
                for (Int is = 1; is <= 4; ++is)
                {
                    std::array<Int, 3> ef = {IN(kIppn1[is - 1][0], parent), IN(kIppn1[is - 1][1], parent),
                                             IN(kIppn1[is - 1][2], parent)};
                    std::sort(ef.begin(), ef.end());
                    if (ef == bf)
                    {
                        matched = is;
                        break;
                    }
                }
                if (matched == 0)
                {
                    throw std::runtime_error("Preprocess::assignBoundaryFaceNumbers - boundary face " +
                                             std::to_string(ib) + " does not match any face of parent tetrahedron " +
                                             std::to_string(parent));
                }
                IS(4, ib) = matched;
            }
        }

        // Preprocess::getNormals (Preprocess.cpp:523-574)
        void get_normals()
        {
            Int bad = 0;
#pragma omp parallel for schedule(static) reduction(max : bad)
            for (Int ie = 1; ie <= nelem; ++ie)
            {
                for (Int is = 1; is <= 4; ++is)
                {
                    const Int p1 = intma[(std::size_t(ie) - 1) * 4 + kIppn1[is - 1][0] - 1];
                    const Int p2 = intma[(std::size_t(ie) - 1) * 4 + kIppn1[is - 1][1] - 1];
                    const Int p3 = intma[(std::size_t(ie) - 1) * 4 + kIppn1[is - 1][2] - 1];
                    const Int popp = intma[(std::size_t(ie) - 1) * 4 + is - 1];
                    const Real* c1 = &coord[(std::size_t(p1) - 1) * 3];
                    const Real* c2 = &coord[(std::size_t(p2) - 1) * 3];
                    const Real* c3 = &coord[(std::size_t(p3) - 1) * 3];
                    const Real* co = &coord[(std::size_t(popp) - 1) * 3];
                    const Real e1x = c2[0] - c1[0], e1y = c2[1] - c1[1], e1z = c2[2] - c1[2];
                    const Real e2x = c3[0] - c1[0], e2y = c3[1] - c1[1], e2z = c3[2] - c1[2];
                    Real ax = e1y * e2z - e1z * e2y;
                    Real ay = e1z * e2x - e1x * e2z;
                    Real az = e1x * e2y - e1y * e2x;
                    const Real amag = std::sqrt(ax * ax + ay * ay + az * az);
                    if (amag <= 0.0 || !std::isfinite(amag))
                    {
                        bad = std::max(bad, 1);
                    }
                    const Real fx = (c1[0] + c2[0] + c3[0]) / 3.0;
                    const Real fy = (c1[1] + c2[1] + c3[1]) / 3.0;
                    const Real fz = (c1[2] + c2[2] + c3[2]) / 3.0;
                    const Real vx = co[0] - fx, vy = co[1] - fy, vz = co[2] - fz;
                    if (ax * vx + ay * vy + az * vz > 0.0)
                    {
                        ax = -ax;
                        ay = -ay;
                        az = -az;
                    }
                    Real* o = &annxf[((std::size_t(ie) - 1) * 4 + is - 1) * 4];
                    o[0] = 0.5 * ax;
                    o[1] = 0.5 * ay;
                    o[2] = 0.5 * az;
                    o[3] = 0.5 * amag;
                }
            }
            if (bad)
            {
                throw std::runtime_error("Preprocess::getNormals - zero-area face");
            }
            for (Int ib = 1; ib <= nboun; ++ib)
            {
                const Int parent = IS(5, ib), lf = IS(4, ib);
                if (parent < 1 || parent > nelem || lf < 1 || lf > 4)
                {
                    throw std::runtime_error(
                        "Preprocess::getNormals - boundary local face numbers must be assigned before boundary normal extraction");
                }
                for (Int c = 1; c <= 4; ++c) FN(c, ib) = ANN(c, lf, parent);
            }
        }

        // Preprocess::massMatrix (Preprocess.cpp:601-679): elcoe_e and the
        // lumped mass; elcoe2/elcoe2p are rebuilt every step by
        // TimeStep::updateLhsDiagonal, so only their inputs are kept.
        void mass_matrix()
        {
            std::fill(Mdiag_real.begin(), Mdiag_real.end(), 0.0);
            std::vector<Real> thermal(std::size_t(npoin), 0.0);
            for (Int ie = 1; ie <= nelem; ++ie)
            {
                if (detJ[ie - 1] <= 0.0 || !std::isfinite(detJ[ie - 1]))
                {
                    throw std::runtime_error("Preprocess::massMatrix - invalid detJ at element " + std::to_string(ie));
                }
                if (rho_cp_e[ie - 1] <= 0.0 || !std::isfinite(rho_cp_e[ie - 1]))
                {
                    throw std::runtime_error("Preprocess::massMatrix - invalid rho*cp at element " + std::to_string(ie));
                }
                const Real nodal_mass = detJ[ie - 1] * (1.0 / 24.0);
                const Real nodal_capacity = rho_cp_e[ie - 1] * nodal_mass;
                for (Int in = 1; in <= 4; ++in)
                {
                    const Int ip = IN(in, ie);
                    Mdiag_real[ip - 1] += nodal_mass;
                    thermal[ip - 1] += nodal_capacity;
                    elcoe_e[(std::size_t(ie) - 1) * 4 + in - 1] = nodal_mass;
                }
            }
            for (Int ip = 1; ip <= npoin; ++ip)
            {
                if (Mdiag_real[ip - 1] <= 0.0 || !std::isfinite(Mdiag_real[ip - 1]))
                {
                    throw std::runtime_error("Preprocess::massMatrix - non-positive lumped mass at node " +
                                             std::to_string(ip));
                }
                if (thermal[ip - 1] <= 0.0 || !std::isfinite(thermal[ip - 1]))
                {
                    throw std::runtime_error("Preprocess::massMatrix - non-positive thermal capacitance at node " +
                                             std::to_string(ip));
                }
            }
        }

        // Preprocess::classifyFaceEdges (Preprocess.cpp:820-843)
        void classify_face_edges()
        {
            std::fill(fedge.begin(), fedge.end(), 0);
            for (Int ib = 1; ib <= nboun; ++ib)
            {
                const Int parent = IS(5, ib), lf = IS(4, ib), bc = IS(6, ib);
                if (parent < 1 || parent > nelem || lf < 1 || lf > 4)
                {
                    throw std::runtime_error("Preprocess::classifyFaceEdges - boundary face number not assigned");
                }
                fedge[(std::size_t(parent) - 1) * 4 + lf - 1] = is_prescribed_velocity_face_bc(bc) ? 1 : 2;
            }
        }

        // Preprocess::elementSize (Preprocess.cpp:863-905)
        void element_size()
        {
            Int bad = 0;
#pragma omp parallel for schedule(static) reduction(max : bad)
            for (Int ie = 1; ie <= nelem; ++ie)
            {
                const Real three_volume = 0.5 * detJ[ie - 1];
                Real hmin = std::numeric_limits<Real>::max();
                for (Int is = 1; is <= 4; ++is)
                {
                    const Real area = annxf[((std::size_t(ie) - 1) * 4 + is - 1) * 4 + 3];
                    if (area <= 0.0 || !std::isfinite(area)) bad = 1;
                    const Real h = three_volume / area;
                    hmin = (h < hmin) ? h : hmin;
                }
                if (hmin <= 0.0 || !std::isfinite(hmin)) bad = 1;
                alen_e[ie - 1] = hmin;
            }
            if (bad) throw std::runtime_error("Preprocess::elementSize - invalid face area or element length");
        }

        // Preprocess::wallDetermination (Preprocess.cpp:932-1031)
        void wall_determination()
        {
            std::vector<Int> node_is_wall(std::size_t(npoin) + 1, 0);
            npoin_wall = 0;
            std::fill(wall_node_list.begin(), wall_node_list.end(), 0);
            std::fill(wall_node_norm.begin(), wall_node_norm.end(), 0.0);
            for (Int ib = 1; ib <= nboun; ++ib)
            {
                const Int bc = IS(6, ib);
                if (!is_no_slip_wall_bc(bc)) continue;
                const Real area = FN(4, ib);
                if (area <= 0.0 || !std::isfinite(area))
                {
                    throw std::runtime_error("Preprocess::wallDetermination - invalid boundary face area");
                }
                for (Int in = 1; in <= 3; ++in)
                {
                    const Int ip = IS(in, ib);
                    if (node_is_wall[ip] == 0)
                    {
                        ++npoin_wall;
                        wall_node_list[npoin_wall - 1] = ip;
                        node_is_wall[ip] = npoin_wall;
                    }
                    const Int iw = node_is_wall[ip];
                    for (Int d = 0; d < 3; ++d)
                    {
                        wall_node_norm[(std::size_t(iw) - 1) * 3 + d] += FN(d + 1, ib);
                    }
                }
            }
            std::vector<char> tf, ts;
            touch_masks(tf, ts);
            for (Int ip = 1; ip <= npoin; ++ip)
            {
                if (!(tf[ip] && ts[ip])) continue;
                if (node_is_wall[ip] == 0)
                {
                    ++npoin_wall;
                    wall_node_list[npoin_wall - 1] = ip;
                    node_is_wall[ip] = npoin_wall;
                }
            }
            for (Int iw = 1; iw <= npoin_wall; ++iw)
            {
                Real* n = &wall_node_norm[(std::size_t(iw) - 1) * 3];
                const Real len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
                if (len > 0.0)
                {
                    n[0] /= len;
                    n[1] /= len;
                    n[2] /= len;
                }
            }
        }

        // Preprocess::computeMassFlowInletVelocity (Preprocess.cpp:1054-1112)
        void mass_flow_inlet_velocity()
        {
            if (cfg.mass_flow_inlet_enabled < 1) return;
            if (cfg.inlet_density <= 0.0 || !std::isfinite(cfg.inlet_density))
            {
                throw std::runtime_error("Preprocess::computeMassFlowInletVelocity - inlet density must be positive");
            }
            Real area_sum = 0.0;
            Int faces = 0;
            for (Int ib = 1; ib <= nboun; ++ib)
            {
                if (IS(6, ib) != 511) continue;
                const Real area = FN(4, ib);
                if (area <= 0.0 || !std::isfinite(area))
                {
                    throw std::runtime_error("Preprocess::computeMassFlowInletVelocity - invalid inlet face area");
                }
                area_sum += area;
                ++faces;
            }
            if (faces < 1)
            {
                throw std::runtime_error(
                    "Preprocess::computeMassFlowInletVelocity - mass-flow inlet enabled but no BC_ID 511 faces were found");
            }
            if (area_sum <= 0.0 || !std::isfinite(area_sum))
            {
                throw std::runtime_error("Preprocess::computeMassFlowInletVelocity - total inlet area is invalid");
            }
            cfg.inlet_u_from_massflow = cfg.inlet_mass_flow_rate / (cfg.inlet_density * area_sum);
        }

        // Preprocess::detectPressureBoundaryNodes (Preprocess.cpp:1152-1242)
        void detect_pressure_boundary_nodes()
        {
            std::vector<char> tf, ts;
            touch_masks(tf, ts);
            std::vector<Int> node_is_fixed(std::size_t(npoin) + 1, 0);
            bc_fixed = 0;
            std::fill(bc_list.begin(), bc_list.end(), 0);
            std::fill(bc_values.begin(), bc_values.end(), 0.0);
            for (Int ib = 1; ib <= nboun; ++ib)
            {
                if (IS(6, ib) != 520) continue;
                for (Int in = 1; in <= 3; ++in)
                {
                    const Int ip = IS(in, ib);
                    if (!tf[ip])
                    {
                        throw std::runtime_error("Preprocess::detectPressureBoundaryNodes - pressure outlet node " +
                                                 std::to_string(ip) + " is not connected to any fluid element");
                    }
                    if (node_is_fixed[ip] != 0) continue;
                    ++bc_fixed;
                    bc_list[bc_fixed - 1] = ip;
                    node_is_fixed[ip] = 1;
                    bc_values[bc_fixed - 1] = cfg.outlet_pressure_gauge;
                }
            }
            if (bc_fixed < 1)
            {
                Int ip = cfg.pnode;
                if (ip < 1 || ip > npoin || !tf[ip])
                {
                    ip = 0;
                    for (Int cand = 1; cand <= npoin; ++cand)
                    {
                        if (tf[cand])
                        {
                            ip = cand;
                            break;
                        }
                    }
                }
                if (ip < 1 || ip > npoin)
                {
                    throw std::runtime_error(
                        "Preprocess::detectPressureBoundaryNodes - no fluid-connected node available for pressure reference");
                }
                bc_fixed = 1;
                bc_list[0] = ip;
                bc_values[0] = cfg.outlet_pressure_gauge;
            }
        }

        // Coloring::build (include/cbs/parallel/Coloring.hpp:27-97): greedy
        // first-fit in element order, counting sort into colour-major order.
        void coloring()
        {
            const std::size_t N = std::size_t(npoin);
            std::vector<Int> n2e_ptr(N + 2, 0), n2e(std::size_t(nelem) * 4), fill(N + 2, 0);
            for (Int ie = 1; ie <= nelem; ++ie)
                for (Int a = 0; a < 4; ++a) ++n2e_ptr[std::size_t(intma[(std::size_t(ie) - 1) * 4 + a]) + 1];
            for (std::size_t ip = 1; ip <= N; ++ip) n2e_ptr[ip + 1] += n2e_ptr[ip];
            for (Int ie = 1; ie <= nelem; ++ie)
                for (Int a = 0; a < 4; ++a)
                {
                    const Int ip = intma[(std::size_t(ie) - 1) * 4 + a];
                    n2e[std::size_t(n2e_ptr[ip] + fill[ip]++)] = ie;
                }
            std::vector<Int> color(std::size_t(nelem) + 1, -1);
            std::vector<Int> stamp;
            ncolor = 0;
            for (Int ie = 1; ie <= nelem; ++ie)
            {
                for (Int a = 0; a < 4; ++a)
                {
                    const Int ip = intma[(std::size_t(ie) - 1) * 4 + a];
                    for (Int k = n2e_ptr[ip]; k < n2e_ptr[std::size_t(ip) + 1]; ++k)
                    {
                        const Int cj = color[std::size_t(n2e[std::size_t(k)])];
                        if (cj >= 0)
                        {
                            if (cj >= Int(stamp.size())) stamp.resize(std::size_t(cj) + 1, -1);
                            stamp[std::size_t(cj)] = ie;
                        }
                    }
                }
                Int c = 0;
                while (c < Int(stamp.size()) && stamp[std::size_t(c)] == ie) ++c;
                color[std::size_t(ie)] = c;
                ncolor = std::max(ncolor, c + 1);
            }
            color_ptr.assign(std::size_t(ncolor) + 1, 0);
            for (Int ie = 1; ie <= nelem; ++ie) ++color_ptr[std::size_t(color[std::size_t(ie)]) + 1];
            for (Int c = 0; c < ncolor; ++c) color_ptr[std::size_t(c) + 1] += color_ptr[std::size_t(c)];
            color_elem.assign(std::size_t(nelem), 0);
            std::vector<Int> off(color_ptr.begin(), color_ptr.end());
            for (Int ie = 1; ie <= nelem; ++ie) color_elem[std::size_t(off[std::size_t(color[std::size_t(ie)])]++)] = ie;
        }

        // wall-clock seconds of each phase of the last setup(), in order
        std::vector<Real> setup_seconds;

        void setup(bool with_coloring = true)
        {
            using Clock = std::chrono::steady_clock;
            setup_seconds.clear();
            auto t = Clock::now();
            auto lap = [&] {
                const auto now = Clock::now();
                setup_seconds.push_back(std::chrono::duration<Real>(now - t).count());
                t = now;
            };
            validate_boundary_flags();
            lap();
            shape_function_derivatives();
            lap();
            assign_boundary_face_numbers();
            lap();
            get_normals();
            lap();
            mass_matrix();
            lap();
            classify_face_edges();
            lap();
            element_size();
            lap();
            wall_determination();
            lap();
            mass_flow_inlet_velocity();
            lap();
            detect_pressure_boundary_nodes();
            lap();
            if (with_coloring)
            {
                coloring();
                lap();
            }
        }
    };

    thread_local std::string g_error;

    template <typename F>
    int guarded(F&& f)
    {
        try
        {
            f();
            return CBS_OK;
        }
        catch (const std::invalid_argument& e)
        {
            g_error = e.what();
            return CBS_ERR_BAD_ARGUMENT;
        }
        catch (const std::exception& e)
        {
            g_error = e.what();
            return CBS_ERR_INVALID_GEOMETRY;
        }
    }

    // ------------------------------------------------------------------
    // Case-file reader (MeshIO::readAll, src/io/MeshIO.cpp:572-1623)
    // ------------------------------------------------------------------
    std::vector<std::string> data_lines(const std::string& path)
    {
        std::ifstream in(path);
        if (!in) throw std::runtime_error("Cannot open file: " + path);
        std::vector<std::string> out;
        std::string raw;
        while (std::getline(in, raw))
        {
            const auto b = raw.find_first_not_of(" \t\r\n");
            if (b == std::string::npos) continue;
            const auto e = raw.find_last_not_of(" \t\r\n");
            std::string t = raw.substr(b, e - b + 1);
            if (t[0] == '#' || t[0] == '!') continue;
            out.push_back(t);
        }
        return out;
    }

    // Positional .par layout: (label, data) pairs (MeshIO.cpp:986-1437).
    void read_par(const std::string& path, cbs_config& c)
    {
        const auto L = data_lines(path);
        std::size_t i = 0;
        auto data = [&]() -> std::istringstream {
            if (i + 1 >= L.size()) throw std::runtime_error("Unexpected end of file: " + path);
            i += 2;
            return std::istringstream(L[i - 1]);
        };
        auto opt_data = [&](bool& ok) -> std::istringstream {
            ok = i + 1 < L.size();
            if (!ok) return std::istringstream("");
            i += 2;
            return std::istringstream(L[i - 1]);
        };
        Int restart = 0, transient = 0, iwrite = 0, write_output = 0, rtmod = 0, dummy_i = 0;
        Real theta = 0, dummy_d = 0;
        data() >> c.solver_opt;
        data() >> restart;
        data() >> c.temp_calc;
        { auto s = data(); s >> c.cinf[0] >> c.cinf[1] >> c.cinf[2] >> c.cinf[3] >> c.cinf[4]; }
        { auto s = data(); s >> c.ntime >> transient >> c.dtfixed >> c.dtfix >> iwrite; c.transient_on = transient; }
        { auto s = data(); s >> c.csafm >> theta; }
        { auto s = data(); s >> c.cbs_scheme >> c.ilots >> c.htype >> c.step2_check >> c.rem_deltp >> c.beta_opt
                           >> c.dtfix_end >> c.csafm2 >> c.epsilon1 >> c.deltr; }
        { auto s = data(); s >> c.re >> c.pr >> dummy_d >> c.ri; }
        data() >> c.convection_type;
        data() >> c.pnode;
        { auto s = data(); s >> write_output; }
        { auto s = data(); s >> c.relToler >> c.absToler; }
        { auto s = data(); s >> dummy_i; }
        { auto s = data(); s >> dummy_i; }
        data() >> rtmod;
        bool ok = false;
        Real alpha_sf = 0, k_ratio = 0, source_solid = 0;
        { auto s = opt_data(ok); if (ok) s >> alpha_sf >> k_ratio >> source_solid >> c.heat_flux_bc; }
        { auto s = opt_data(ok); if (ok) s >> c.dimensional_mode >> c.material_properties_enabled >> c.mass_flow_inlet_enabled; }
        { auto s = opt_data(ok); Real pref = 0, depth = 0;
          if (ok) s >> pref >> depth >> c.inlet_mass_flow_rate >> c.inlet_density >> c.inlet_temperature >> c.outlet_pressure_gauge; }
        { auto s = opt_data(ok); if (ok) s >> c.inlet_u >> c.inlet_v >> c.inlet_w; }
        { auto s = opt_data(ok); }
        { auto s = data(); Real tol = 0; s >> c.cg_preconditioner >> c.cg_conv_test >> c.cg_max_iter >> tol; }
    }
}

using cbs_host::Case;
using cbs_host::Int;
using cbs_host::Real;
using cbs_host::guarded;

struct cbs_case
{
    Case c;
};

extern "C" const char* cbs_case_last_error(void) { return cbs_host::g_error.c_str(); }

static void default_config(cbs_config& c)
{
    std::memset(&c, 0, sizeof c);
    c.solver_opt = 1;
    c.cbs_scheme = 1;
    c.htype = 1;
    c.beta_opt = 1;
    c.dtfix = 1.0;
    c.csafm = 0.5;
    c.deltr = 1.0e20;
    c.csafm2 = 2.0;
    c.epsilon1 = 1.0e-6;
    c.cg_preconditioner = 1;
    c.cg_conv_test = 3;
    c.cg_max_iter = 100000;
    c.inlet_density = 1.0;
    c.inlet_temperature = 568.15;
    c.pnode = 1;
}

extern "C" int cbs_case_create(const cbs_config* par, int32_t nelem, int32_t npoin, int32_t nboun,
                               const double* coord, const int32_t* intma, const int32_t* bface,
                               const int32_t* mat_elem, int32_t nmat, const cbs_material* mats,
                               cbs_case** out)
{
    if (!par || !coord || !intma || (nboun > 0 && !bface) || !out || nelem < 1 || npoin < 1 || nboun < 0)
    {
        cbs_host::g_error = "cbs_case_create - invalid argument";
        return CBS_ERR_BAD_ARGUMENT;
    }
    auto* h = new cbs_case();
    const int rc = guarded([&] {
        Case& c = h->c;
        c.cfg = *par;
        c.nelem = nelem;
        c.npoin = npoin;
        c.nboun = nboun;
        c.allocate();
        std::memcpy(c.coord.data(), coord, sizeof(double) * 3 * std::size_t(npoin));
        for (std::size_t i = 0; i < 4 * std::size_t(nelem); ++i)
        {
            if (intma[i] < 1 || intma[i] > npoin)
            {
                throw std::invalid_argument("MeshIO::readMeshFile - connectivity node out of range");
            }
            c.intma[i] = intma[i];
        }
        for (Int ib = 1; ib <= nboun; ++ib)
        {
            const int32_t* r = bface + (std::size_t(ib) - 1) * 5;
            for (int k = 0; k < 3; ++k)
            {
                if (r[k] < 1 || r[k] > npoin) throw std::invalid_argument("boundary-face node out of range");
            }
            if (r[3] < 1 || r[3] > nelem) throw std::invalid_argument("boundary-face parent element out of range");
            c.IS(1, ib) = r[0];
            c.IS(2, ib) = r[1];
            c.IS(3, ib) = r[2];
            c.IS(4, ib) = 0;
            c.IS(5, ib) = r[3];
            c.IS(6, ib) = r[4];
        }
        c.initialise_fields(mat_elem, nmat, mats);
    });
    if (rc != CBS_OK)
    {
        delete h;
        return rc;
    }
    *out = h;
    return CBS_OK;
}

namespace cbs_host
{
    // Binary mesh: the content of <base>.plt (and <base>.material) as raw
    // little-endian arrays.
    //   char magic[8] = "CBSMESH1"; uint32 version = 1; uint32 flags (bit 0:
    //   material ids present); int32 nelem, npoin, nboun, 0;
    //   double coord[3 npoin] (node-major); int32 intma[4 nelem] (1-based);
    //   int32 bface[5 nboun] (n1 n2 n3 parent raw_side_id);
    //   [int32 mat[nelem]];
    //   uint64 checksum of the arrays (sum of their 64-bit words, each times
    //   an odd multiplier depending on its position; an odd trailing 4-byte
    //   word is padded with zeros).
    constexpr char kMeshMagic[8] = {'C', 'B', 'S', 'M', 'E', 'S', 'H', '1'};

    struct MeshChecksum
    {
        uint64_t sum = 0, pos = 0;
        uint32_t carry = 0;
        bool odd = false;
        void word(uint64_t w)
        {
            sum += w * (2 * pos + 1);
            ++pos;
        }
        void add(const void* data, size_t bytes)
        {
            const auto* p = static_cast<const unsigned char*>(data);
            size_t k = 0;
            if (odd && bytes >= 4)
            {
                uint32_t hi;
                std::memcpy(&hi, p, 4);
                word(uint64_t(carry) | (uint64_t(hi) << 32));
                odd = false;
                k = 4;
            }
            for (; k + 8 <= bytes; k += 8)
            {
                uint64_t w;
                std::memcpy(&w, p + k, 8);
                word(w);
            }
            if (k + 4 == bytes)
            {
                std::memcpy(&carry, p + k, 4);
                odd = true;
            }
        }
        uint64_t finish()
        {
            if (odd) word(uint64_t(carry));
            odd = false;
            return sum;
        }
    };

    // MeshIO::readMeshFile (src/io/MeshIO.cpp:713-868): sizes, then element,
    // node and boundary-face records of <base>.plt.
    void read_mesh_ascii(const std::string& base, Int& nelem, Int& npoin, Int& nboun, std::vector<double>& coord,
                         std::vector<Int>& intma, std::vector<Int>& bface)
    {
        std::ifstream in(base + ".plt");
        if (!in) throw std::runtime_error("MeshIO::readSizes - cannot open mesh file: " + base + ".plt");
        in >> nelem >> npoin >> nboun;
        if (!in || nelem < 1 || npoin < 1 || nboun < 0)
            throw std::runtime_error("MeshIO::readSizes - non-positive mesh size in .plt header");
        intma.assign(4 * std::size_t(nelem), 0);
        coord.assign(3 * std::size_t(npoin), 0.0);
        bface.assign(5 * std::size_t(nboun), 0);
        for (Int i = 0; i < nelem; ++i)
        {
            Int ie = 0;
            in >> ie;
            if (ie < 1 || ie > nelem) throw std::runtime_error("MeshIO::readMeshFile - element index out of range");
            for (int a = 0; a < 4; ++a) in >> intma[(std::size_t(ie) - 1) * 4 + a];
        }
        for (Int i = 0; i < npoin; ++i)
        {
            Int ip = 0;
            in >> ip;
            if (ip < 1 || ip > npoin) throw std::runtime_error("MeshIO::readMeshFile - node index out of range");
            for (int d = 0; d < 3; ++d) in >> coord[(std::size_t(ip) - 1) * 3 + d];
        }
        for (Int ib = 0; ib < nboun; ++ib)
            for (int k = 0; k < 5; ++k) in >> bface[std::size_t(ib) * 5 + k];
        if (!in) throw std::runtime_error("MeshIO::readMeshFile - truncated mesh file");
    }

    // MeshIO::readMaterialFile: element material ids of <base>.material.
    std::vector<Int> read_material_ids(const std::string& base, Int nelem)
    {
        const auto ml = data_lines(base + ".material");
        std::size_t first = 0;
        {
            std::istringstream s(ml.at(0));
            std::vector<Int> v;
            Int x;
            while (s >> x) v.push_back(x);
            if (v.size() == 1) first = 1;
        }
        std::vector<Int> mat(std::size_t(nelem), 0);
        for (Int i = 0; i < nelem; ++i)
        {
            std::istringstream s(ml.at(first + std::size_t(i)));
            Int ie, n1, n2, n3, n4, m;
            s >> ie >> n1 >> n2 >> n3 >> n4 >> m;
            if (ie < 1 || ie > nelem) throw std::runtime_error("MeshIO::readMaterialFile - element index out of range");
            mat[std::size_t(ie) - 1] = m;
        }
        return mat;
    }

    bool file_exists(const std::string& path)
    {
        std::ifstream f(path, std::ios::binary);
        return bool(f);
    }

    // Modification time of a file in nanoseconds, or -1 when it is absent.
    int64_t mtime_ns(const std::string& path)
    {
        struct stat st;
        if (::stat(path.c_str(), &st) != 0) return -1;
        return int64_t(st.st_mtim.tv_sec) * 1000000000 + int64_t(st.st_mtim.tv_nsec);
    }

    // A binary mesh converted from text files that were edited after it
    // would set up the old mesh: refuse it.
    void check_binary_mesh_current(const std::string& base)
    {
        const int64_t bin = mtime_ns(base + ".cbsmesh");
        for (const char* ext : {".plt", ".material"})
        {
            const int64_t t = mtime_ns(base + ext);
            if (t > bin)
                throw std::runtime_error("MeshIO::readMeshFile - binary mesh " + base + ".cbsmesh is older than " +
                                         base + ext + "; remove it or convert again");
        }
    }

    void write_mesh_binary(const std::string& path, Int nelem, Int npoin, Int nboun, const double* coord,
                           const Int* intma, const Int* bface, const Int* mat)
    {
        std::ofstream f(path + ".part", std::ios::binary | std::ios::trunc);
        if (!f) throw std::runtime_error("cbs_case_write_mesh_binary - cannot open " + path);
        const uint32_t version = 1, flags = mat ? 1u : 0u;
        const int32_t sizes[4] = {nelem, npoin, nboun, 0};
        f.write(kMeshMagic, 8);
        f.write(reinterpret_cast<const char*>(&version), 4);
        f.write(reinterpret_cast<const char*>(&flags), 4);
        f.write(reinterpret_cast<const char*>(sizes), sizeof sizes);
        MeshChecksum ck;
        auto put = [&](const void* p, size_t bytes) {
            f.write(static_cast<const char*>(p), std::streamsize(bytes));
            ck.add(p, bytes);
        };
        put(coord, sizeof(double) * 3 * std::size_t(npoin));
        put(intma, sizeof(Int) * 4 * std::size_t(nelem));
        put(bface, sizeof(Int) * 5 * std::size_t(nboun));
        if (mat) put(mat, sizeof(Int) * std::size_t(nelem));
        const uint64_t sum = ck.finish();
        f.write(reinterpret_cast<const char*>(&sum), 8);
        f.close();
        if (!f) throw std::runtime_error("cbs_case_write_mesh_binary - cannot write " + path);
        if (std::rename((path + ".part").c_str(), path.c_str()) != 0)
            throw std::runtime_error("cbs_case_write_mesh_binary - cannot rename to " + path);
    }

    // Reads a binary mesh; mat is filled only when the file holds material ids.
    void read_mesh_binary(const std::string& path, Int& nelem, Int& npoin, Int& nboun, std::vector<double>& coord,
                          std::vector<Int>& intma, std::vector<Int>& bface, std::vector<Int>& mat)
    {
        std::ifstream f(path, std::ios::binary);
        if (!f) throw std::runtime_error("MeshIO::readMeshFile - cannot open binary mesh " + path);
        char magic[8];
        uint32_t version = 0, flags = 0;
        int32_t sizes[4];
        f.read(magic, 8);
        f.read(reinterpret_cast<char*>(&version), 4);
        f.read(reinterpret_cast<char*>(&flags), 4);
        f.read(reinterpret_cast<char*>(sizes), sizeof sizes);
        if (!f || std::memcmp(magic, kMeshMagic, 8) != 0 || version != 1)
            throw std::runtime_error("MeshIO::readMeshFile - not a version-1 binary mesh: " + path);
        nelem = sizes[0];
        npoin = sizes[1];
        nboun = sizes[2];
        if (nelem < 1 || npoin < 1 || nboun < 0)
            throw std::runtime_error("MeshIO::readSizes - non-positive mesh size in binary mesh header");
        MeshChecksum ck;
        auto get = [&](auto& v, size_t n) {
            v.resize(n);
            const size_t bytes = n * sizeof(v[0]);
            f.read(reinterpret_cast<char*>(v.data()), std::streamsize(bytes));
            if (!f) throw std::runtime_error("MeshIO::readMeshFile - truncated binary mesh " + path);
            ck.add(v.data(), bytes);
        };
        get(coord, 3 * std::size_t(npoin));
        get(intma, 4 * std::size_t(nelem));
        get(bface, 5 * std::size_t(nboun));
        if (flags & 1u) get(mat, std::size_t(nelem));
        uint64_t sum = 0;
        f.read(reinterpret_cast<char*>(&sum), 8);
        if (!f || sum != ck.finish())
            throw std::runtime_error("MeshIO::readMeshFile - binary mesh checksum mismatch: " + path);
    }
}

extern "C" int cbs_case_read(const char* base_c, cbs_case** out)
{
    if (!base_c || !out)
    {
        cbs_host::g_error = "cbs_case_read - invalid argument";
        return CBS_ERR_BAD_ARGUMENT;
    }
    const std::string base(base_c);
    cbs_config cfg;
    default_config(cfg);
    std::vector<double> coord;
    std::vector<int32_t> intma, bface, mat;
    std::vector<cbs_material> mats;
    Int nelem = 0, npoin = 0, nboun = 0;
    int rc = guarded([&] {
        cbs_host::read_par(base + ".par", cfg);
        std::vector<Int> bin_mat;
        const bool binary = cbs_host::file_exists(base + ".cbsmesh");
        if (binary)
        {
            cbs_host::check_binary_mesh_current(base);
            cbs_host::read_mesh_binary(base + ".cbsmesh", nelem, npoin, nboun, coord, intma, bface, bin_mat);
        }
        else
        {
            cbs_host::read_mesh_ascii(base, nelem, npoin, nboun, coord, intma, bface);
        }
        // .bco: raw side id -> solver BC (MeshIO.cpp:872-982)
        const auto bl = cbs_host::data_lines(base + ".bco");
        if (bl.empty()) throw std::runtime_error("MeshIO::readBoundaryFile - invalid .bco header");
        Int nflag = 0, mass_check = 0;
        { std::istringstream s(bl[0]); s >> nflag; if (!(s >> mass_check)) mass_check = 0; }
        std::unordered_map<Int, Int> map;
        for (Int i = 1; i <= nflag; ++i)
        {
            if (std::size_t(i) >= bl.size()) throw std::runtime_error("Unexpected end of file: " + base + ".bco");
            std::istringstream s(bl[std::size_t(i)]);
            Int raw = 0, bc = 0;
            s >> raw >> bc;
            if (bc == 504 || !cbs_host::supported_bc(bc))
                throw std::runtime_error("MeshIO::readBoundaryFile - unsupported solver BC ID " + std::to_string(bc));
            if (!map.emplace(raw, bc).second)
                throw std::runtime_error("MeshIO::readBoundaryFile - duplicate raw boundary-side ID in .bco");
        }
        for (Int ib = 0; ib < nboun; ++ib)
        {
            const auto it = map.find(bface[std::size_t(ib) * 5 + 4]);
            if (it == map.end()) throw std::runtime_error("MeshIO::readBoundaryFile - boundary-face raw ID not found in .bco file");
            bface[std::size_t(ib) * 5 + 4] = it->second;
        }
        if (cfg.material_properties_enabled > 0)
        {
            if (!bin_mat.empty())
            {
                mat = bin_mat;
            }
            else
            {
                mat = cbs_host::read_material_ids(base, nelem);
            }
            std::ifstream mp(base + ".matprop");
            Int nmat = 0;
            mp >> nmat;
            for (Int i = 0; i < nmat; ++i)
            {
                cbs_material m{};
                std::string name, phase;
                mp >> m.id >> name >> phase >> m.rho >> m.cp >> m.k >> m.mu >> m.qvol;
                if (!mp) throw std::runtime_error("MeshIO::readMaterialPropertyFile - invalid .matprop record");
                mats.push_back(m);
            }
        }
    });
    if (rc != CBS_OK) return rc;
    return cbs_case_create(&cfg, nelem, npoin, nboun, coord.data(), intma.data(), bface.data(),
                           mat.empty() ? nullptr : mat.data(), Int(mats.size()),
                           mats.empty() ? nullptr : mats.data(), out);
}

extern "C" int cbs_case_write_mesh_binary(const char* base_c)
{
    if (!base_c)
    {
        cbs_host::g_error = "cbs_case_write_mesh_binary - invalid argument";
        return CBS_ERR_BAD_ARGUMENT;
    }
    const std::string base(base_c);
    if (cbs_host::file_exists(base + ".cbsmesh"))
    {
        cbs_host::g_error = "cbs_case_write_mesh_binary - " + base + ".cbsmesh exists";
        return CBS_ERR_BAD_ARGUMENT;
    }
    return guarded([&] {
        Int nelem = 0, npoin = 0, nboun = 0;
        std::vector<double> coord;
        std::vector<Int> intma, bface, mat;
        cbs_host::read_mesh_ascii(base, nelem, npoin, nboun, coord, intma, bface);
        if (cbs_host::file_exists(base + ".material")) mat = cbs_host::read_material_ids(base, nelem);
        cbs_host::write_mesh_binary(base + ".cbsmesh", nelem, npoin, nboun, coord.data(), intma.data(), bface.data(),
                                    mat.empty() ? nullptr : mat.data());
    });
}

extern "C" int cbs_case_write_mesh_arrays_binary(const char* path, int32_t nelem, int32_t npoin, int32_t nboun,
                                                 const double* coord, const int32_t* intma, const int32_t* bface,
                                                 const int32_t* mat_elem)
{
    if (!path || nelem < 1 || npoin < 1 || nboun < 0 || !coord || !intma || (nboun > 0 && !bface))
    {
        cbs_host::g_error = "cbs_case_write_mesh_arrays_binary - invalid argument";
        return CBS_ERR_BAD_ARGUMENT;
    }
    return guarded([&] { cbs_host::write_mesh_binary(path, nelem, npoin, nboun, coord, intma, bface, mat_elem); });
}

extern "C" int cbs_case_setup(cbs_case* h)
{
    if (!h)
    {
        cbs_host::g_error = "cbs_case_setup - null case";
        return CBS_ERR_BAD_ARGUMENT;
    }
    return guarded([&] { h->c.setup(); });
}

extern "C" int cbs_case_setup_without_coloring(cbs_case* h)
{
    if (!h)
    {
        cbs_host::g_error = "cbs_case_setup_without_coloring - null case";
        return CBS_ERR_BAD_ARGUMENT;
    }
    return guarded([&] {
        h->c.setup(false);
        h->c.ncolor = 0;
        h->c.color_ptr.assign(1, 0);
        h->c.color_elem.clear();
    });
}

extern "C" int cbs_case_set_coloring(cbs_case* h, int32_t ncolor, const int32_t* color_ptr, const int32_t* color_elem)
{
    if (!h || ncolor < 1 || !color_ptr || !color_elem)
    {
        cbs_host::g_error = "cbs_case_set_coloring - invalid argument";
        return CBS_ERR_BAD_ARGUMENT;
    }
    return guarded([&] {
        Case& c = h->c;
        const std::size_t E = std::size_t(c.nelem);
        if (color_ptr[0] != 0 || color_ptr[ncolor] != c.nelem)
            throw std::invalid_argument("cbs_case_set_coloring - colour offsets do not cover the elements");
        for (Int k = 0; k < ncolor; ++k)
            if (color_ptr[k + 1] <= color_ptr[k])
                throw std::invalid_argument("cbs_case_set_coloring - empty or decreasing colour class");
        std::vector<char> seen(E + 1, 0);
        for (std::size_t i = 0; i < E; ++i)
        {
            const Int e = color_elem[i];
            if (e < 1 || e > c.nelem || seen[std::size_t(e)]++)
                throw std::invalid_argument("cbs_case_set_coloring - colour classes are not a partition of the elements");
        }
        c.ncolor = ncolor;
        c.color_ptr.assign(color_ptr, color_ptr + ncolor + 1);
        c.color_elem.assign(color_elem, color_elem + E);
    });
}

extern "C" int cbs_case_config(const cbs_case* h, cbs_config* out)
{
    if (!h || !out) return CBS_ERR_BAD_ARGUMENT;
    *out = h->c.cfg;
    return CBS_OK;
}

extern "C" int cbs_case_mesh_desc(const cbs_case* h, cbs_mesh_desc* d)
{
    if (!h || !d) return CBS_ERR_BAD_ARGUMENT;
    const Case& c = h->c;
    std::memset(d, 0, sizeof *d);
    d->npoin = c.npoin;
    d->nelem = c.nelem;
    d->nboun = c.nboun;
    d->ncolor = c.ncolor;
    d->npoin_wall = c.npoin_wall;
    d->bc_fixed = c.bc_fixed;
    d->intma = c.intma.data();
    d->coord = c.coord.data();
    d->mat_elem = c.mat_elem.data();
    d->rho_e = c.rho_e.data();
    d->mu_e = c.mu_e.data();
    d->k_e = c.k_e.data();
    d->rho_cp_e = c.rho_cp_e.data();
    d->alpha_e = c.alpha_e.data();
    d->Qvol_e = c.Qvol_e.data();
    d->dNkdx = c.dNkdx.data();
    d->detJ = c.detJ.data();
    d->annxf = c.annxf.data();
    d->fedge = c.fedge.data();
    d->elcoe_e = c.elcoe_e.data();
    d->alen_e = c.alen_e.data();
    d->iside = c.iside.data();
    d->face_norm = c.face_norm.data();
    d->wall_node_list = c.wall_node_list.data();
    d->bc_list = c.bc_list.data();
    d->bc_values = c.bc_values.data();
    d->color_ptr = c.color_ptr.data();
    d->color_elem = c.color_elem.data();
    return CBS_OK;
}

extern "C" int cbs_case_array(const cbs_case* h, const char* name, const void** data, int64_t* count,
                              char* dtype)
{
    if (!h || !name || !data || !count || !dtype) return CBS_ERR_BAD_ARGUMENT;
    const Case& c = h->c;
    const std::string n(name);
    auto put_d = [&](const std::vector<double>& v, std::size_t cnt) {
        *data = v.data();
        *count = int64_t(cnt);
        *dtype = 'd';
    };
    auto put_i = [&](const std::vector<int32_t>& v, std::size_t cnt) {
        *data = v.data();
        *count = int64_t(cnt);
        *dtype = 'i';
    };
    if (n == "setup_seconds") put_d(c.setup_seconds, c.setup_seconds.size());
    else if (n == "intma") put_i(c.intma, c.intma.size());
    else if (n == "iside") put_i(c.iside, c.iside.size());
    else if (n == "mat_elem") put_i(c.mat_elem, c.mat_elem.size());
    else if (n == "fedge") put_i(c.fedge, c.fedge.size());
    else if (n == "wall_node_list") put_i(c.wall_node_list, std::size_t(c.npoin_wall));
    else if (n == "bc_list") put_i(c.bc_list, std::size_t(c.bc_fixed));
    else if (n == "color_ptr") put_i(c.color_ptr, c.color_ptr.size());
    else if (n == "color_elem") put_i(c.color_elem, c.color_elem.size());
    else if (n == "coord") put_d(c.coord, c.coord.size());
    else if (n == "unkno") put_d(c.unkno, c.unkno.size());
    else if (n == "pres") put_d(c.pres, c.pres.size());
    else if (n == "temperature") put_d(c.temperature, c.temperature.size());
    else if (n == "rho_e") put_d(c.rho_e, c.rho_e.size());
    else if (n == "cp_e") put_d(c.cp_e, c.cp_e.size());
    else if (n == "k_e") put_d(c.k_e, c.k_e.size());
    else if (n == "mu_e") put_d(c.mu_e, c.mu_e.size());
    else if (n == "rho_cp_e") put_d(c.rho_cp_e, c.rho_cp_e.size());
    else if (n == "alpha_e") put_d(c.alpha_e, c.alpha_e.size());
    else if (n == "Qvol_e") put_d(c.Qvol_e, c.Qvol_e.size());
    else if (n == "dNkdx") put_d(c.dNkdx, c.dNkdx.size());
    else if (n == "detJ") put_d(c.detJ, c.detJ.size());
    else if (n == "annxf") put_d(c.annxf, c.annxf.size());
    else if (n == "face_norm") put_d(c.face_norm, c.face_norm.size());
    else if (n == "elcoe_e") put_d(c.elcoe_e, c.elcoe_e.size());
    else if (n == "alen_e") put_d(c.alen_e, c.alen_e.size());
    else if (n == "Mdiag_real") put_d(c.Mdiag_real, c.Mdiag_real.size());
    else if (n == "bc_values") put_d(c.bc_values, std::size_t(c.bc_fixed));
    else
    {
        cbs_host::g_error = "cbs_case_array - unknown array " + n;
        return CBS_ERR_BAD_ARGUMENT;
    }
    return CBS_OK;
}

extern "C" void cbs_case_destroy(cbs_case* h) { delete h; }

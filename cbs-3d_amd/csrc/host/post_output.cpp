// Solution output of a CBS run: residual CSV, VTU and PVD files with the
// reference's names, columns and arrays (src/io/Post.cpp), written off the
// time step's critical path.
//
// What is alike by necessity: the directory layout and file names
// (Post.cpp:61-108), the CSV header and row text (:534-558, 768-812), the PVD
// text (:1133-1182), the VTU arrays with their names, types, order and masks
// (:436-514, 887-1130) and the output schedule (:571-604, 731-754, 821-860).
// A caller that switches to this writer keeps every file a post-processing
// script reads.  What differs: VTU arrays are stored as one appended
// raw-binary block (VTK XML "appended", UInt64 headers) instead of ASCII, the
// static part of every VTU (points, cells, node masks, cell data) is encoded
// once per run, and VTU files can be written by a background thread from
// pinned snapshots of the device fields.
#include "cbs_post.h"

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <filesystem>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace
{
    namespace fs = std::filesystem;

    thread_local std::string g_last_error;

    struct PostFailure : std::runtime_error
    {
        int status;
        PostFailure(int s, const std::string& m) : std::runtime_error(m), status(s) {}
    };

    // Post.cpp:117-122: NaN and infinities are written as zero.
    inline double safe_value(double v) { return std::isfinite(v) ? v : 0.0; }

    // The reference streams doubles with setprecision(16) in the default
    // float field, which is printf's %.16g.
    inline void put_g16(std::string& out, double v)
    {
        char buf[40];
        const int n = std::snprintf(buf, sizeof buf, "%.16g", v);
        out.append(buf, size_t(n));
    }

    std::string step_tag(int32_t iitime)
    {
        char buf[32];
        std::snprintf(buf, sizeof buf, "step_%08d", iitime);
        return buf;
    }

    // One array of the appended block: its XML element and its bytes.
    struct ArrayRef
    {
        std::string type, name;
        int ncomp = 1;
        uint64_t bytes = 0;
        uint64_t offset = 0;  // of its UInt64 size header inside the block
    };

    void write_all(int fd, const void* data, size_t n, const std::string& path)
    {
        const char* p = static_cast<const char*>(data);
        while (n > 0)
        {
            const ssize_t w = ::write(fd, p, n);
            if (w < 0)
            {
                if (errno == EINTR) continue;
                throw PostFailure(CBS_ERR_IO, "Post::writeVTU - cannot write VTU file " + path);
            }
            p += w;
            n -= size_t(w);
        }
    }

    struct Job
    {
        int32_t iitime = 0;
        cbs_post_fields fields{};
        cbs_post_ready_fn ready = nullptr;
        cbs_post_done_fn done = nullptr;
        void* user = nullptr;
        std::vector<int32_t> hist_it;  // output history up to this job
        std::vector<double> hist_t;
    };
}  // namespace

struct cbs_post
{
    cbs_post_config cfg{};
    int32_t N = 0, E = 0, B = 0;
    fs::path dir;
    std::string case_name;

    // node masks (Post.cpp:436-514)
    std::vector<uint8_t> pressure_active, velocity_active;

    // VTU: XML before the block, the static block, the XML after it; the
    // dynamic arrays (pressure, temperature, velocity, velocity_magnitude)
    // open the block
    std::string head, tail;
    std::vector<char> static_block;
    uint64_t dynamic_bytes = 0;

    // residual log and output schedule
    FILE* csv = nullptr;
    std::vector<int32_t> out_it;
    std::vector<double> out_t;
    double next_vtu_time = 0.0;

    // writer thread
    std::mutex mu;
    std::condition_variable cv_work, cv_idle;
    std::deque<Job> queue;
    bool busy = false, stop = false;
    std::thread worker;
    int bg_status = CBS_OK;
    std::string bg_message;

    // statistics
    int64_t vtu_files = 0, vtu_bytes = 0;
    double write_seconds = 0.0;
    std::vector<char> scratch;  // dynamic arrays of one file (writer side)

    fs::path residual_path() const { return dir / (case_name + "_residuals.csv"); }
    fs::path pvd_path() const { return dir / (case_name + ".pvd"); }
    std::string vtu_name(int32_t it) const { return case_name + "_" + step_tag(it) + ".vtu"; }
};

namespace
{
    template <typename F>
    int guarded(F&& f)
    {
        try
        {
            f();
            return CBS_OK;
        }
        catch (const PostFailure& e)
        {
            g_last_error = e.what();
            return e.status;
        }
        catch (const std::exception& e)
        {
            g_last_error = e.what();
            return CBS_ERR_IO;
        }
    }

    template <typename T>
    void append_array(std::vector<char>& blk, const std::vector<T>& v)
    {
        const uint64_t n = uint64_t(v.size()) * sizeof(T);
        const size_t at = blk.size();
        blk.resize(at + sizeof n + n);
        std::memcpy(blk.data() + at, &n, sizeof n);
        if (n) std::memcpy(blk.data() + at + sizeof n, v.data(), n);
    }

    std::string data_array_xml(const ArrayRef& a, const char* indent)
    {
        std::string s = indent;
        s += "<DataArray type=\"" + a.type + "\"";
        if (!a.name.empty()) s += " Name=\"" + a.name + "\"";
        s += " NumberOfComponents=\"" + std::to_string(a.ncomp) + "\" format=\"appended\" offset=\"" +
             std::to_string(a.offset) + "\"/>\n";
        return s;
    }

    // Builds the static block and the XML around the appended data.
    void build_static(cbs_post* p, const cbs_post_mesh& m)
    {
        const int32_t N = m.npoin, E = m.nelem, B = m.nboun;
        std::vector<uint8_t> touches_fluid(static_cast<size_t>(N), 0), touches_solid(static_cast<size_t>(N), 0);
        for (int32_t e = 0; e < E; ++e)
        {
            const bool fluid = m.mat_elem[e] == 0;
            for (int a = 0; a < 4; ++a)
            {
                const int32_t ip = m.intma[size_t(e) * 4 + size_t(a)];
                if (ip < 1 || ip > N)
                    throw PostFailure(CBS_ERR_BAD_ARGUMENT, "Post::writeVTU - element connectivity node out of range");
                (fluid ? touches_fluid : touches_solid)[size_t(ip - 1)] = 1;
            }
        }
        const size_t nn = static_cast<size_t>(N);
        std::vector<int32_t> domain(nn), pact(nn), vact(nn), tact(nn);
        p->pressure_active.assign(size_t(N), 0);
        p->velocity_active.assign(size_t(N), 0);
        for (int32_t i = 0; i < N; ++i)
        {
            const bool f = touches_fluid[size_t(i)] != 0, s = touches_solid[size_t(i)] != 0;
            pact[size_t(i)] = f ? 1 : 0;
            vact[size_t(i)] = (f && !s) ? 1 : 0;
            tact[size_t(i)] = (f || s) ? 1 : 0;
            domain[size_t(i)] = (f && s) ? 3 : f ? 1 : s ? 2 : 0;
            p->pressure_active[size_t(i)] = uint8_t(pact[size_t(i)]);
            p->velocity_active[size_t(i)] = uint8_t(vact[size_t(i)]);
        }

        const size_t C = size_t(E) + size_t(B);
        std::vector<double> points(m.coord, m.coord + size_t(N) * 3);
        std::vector<int32_t> conn;
        conn.reserve(size_t(E) * 4 + size_t(B) * 3);
        for (int32_t e = 0; e < E; ++e)
            for (int a = 0; a < 4; ++a) conn.push_back(m.intma[size_t(e) * 4 + size_t(a)] - 1);
        for (int32_t b = 0; b < B; ++b)
            for (int a = 0; a < 3; ++a) conn.push_back(m.iside[size_t(b) * 6 + size_t(a)] - 1);
        std::vector<int32_t> offsets(C);
        int32_t off = 0;
        for (size_t c = 0; c < C; ++c) offsets[c] = (off += (c < size_t(E) ? 4 : 3));
        std::vector<uint8_t> types(C);
        for (size_t c = 0; c < C; ++c) types[c] = c < size_t(E) ? 10 : 5;
        std::vector<int32_t> kind(C), mat(C), bc(C), parent(C);
        for (int32_t e = 0; e < E; ++e)
        {
            kind[size_t(e)] = 1;
            mat[size_t(e)] = m.mat_elem[e];
            bc[size_t(e)] = 0;
            parent[size_t(e)] = e + 1;
        }
        for (int32_t b = 0; b < B; ++b)
        {
            const size_t c = size_t(E) + size_t(b);
            kind[c] = 2;
            mat[c] = -1;
            bc[c] = m.iside[size_t(b) * 6 + 5];      // bsid = 6
            parent[c] = m.iside[size_t(b) * 6 + 4];  // nsidpe = 5
        }
        // Post.cpp:408-428: the last face visited sets a node's id
        std::vector<int32_t> node_bc;
        if (p->cfg.write_boundary_debug_arrays > 0)
        {
            node_bc.assign(size_t(N), 0);
            for (int32_t b = 0; b < B; ++b)
            {
                const int32_t id = m.iside[size_t(b) * 6 + 5];
                for (int a = 0; a < 3; ++a)
                {
                    const int32_t ip = m.iside[size_t(b) * 6 + size_t(a)];
                    if (ip >= 1 && ip <= N) node_bc[size_t(ip - 1)] = id;
                }
            }
        }

        // block layout: dynamic arrays first (fixed sizes), then the static
        ArrayRef pres{"Float64", "pressure", 1}, temp{"Float64", "temperature", 1}, vel{"Float64", "velocity", 3},
            vmag{"Float64", "velocity_magnitude", 1};
        uint64_t at = 0;
        for (ArrayRef* a : {&pres, &temp, &vel, &vmag})
        {
            a->bytes = uint64_t(N) * uint64_t(a->ncomp) * 8;
            a->offset = at;
            at += 8 + a->bytes;
        }
        p->dynamic_bytes = at;
        std::vector<char>& blk = p->static_block;
        blk.clear();
        auto place = [&](ArrayRef r, auto const& v) {
            r.offset = at + blk.size();
            append_array(blk, v);
            return r;
        };
        const ArrayRef r_points = place({"Float64", "", 3}, points);
        const ArrayRef r_conn = place({"Int32", "connectivity", 1}, conn);
        const ArrayRef r_offs = place({"Int32", "offsets", 1}, offsets);
        const ArrayRef r_types = place({"UInt8", "types", 1}, types);
        const ArrayRef r_domain = place({"Int32", "node_domain_kind", 1}, domain);
        const ArrayRef r_pact = place({"Int32", "pressure_active", 1}, pact);
        const ArrayRef r_vact = place({"Int32", "velocity_active", 1}, vact);
        const ArrayRef r_tact = place({"Int32", "thermal_active", 1}, tact);
        ArrayRef r_nbc;
        if (p->cfg.write_boundary_debug_arrays > 0) r_nbc = place({"Int32", "node_bc_id_debug", 1}, node_bc);
        const ArrayRef r_kind = place({"Int32", "cell_kind", 1}, kind);
        const ArrayRef r_mat = place({"Int32", "mat_id", 1}, mat);
        const ArrayRef r_bc = place({"Int32", "bc_id", 1}, bc);
        const ArrayRef r_parent = place({"Int32", "parent_element", 1}, parent);

        std::string h;
        h += "<?xml version=\"1.0\"?>\n";
        h += "<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n";
        h += "  <UnstructuredGrid>\n";
        h += "    <Piece NumberOfPoints=\"" + std::to_string(N) + "\" NumberOfCells=\"" + std::to_string(C) + "\">\n";
        h += "      <Points>\n" + data_array_xml(r_points, "        ") + "      </Points>\n";
        h += "      <Cells>\n" + data_array_xml(r_conn, "        ") + data_array_xml(r_offs, "        ") +
             data_array_xml(r_types, "        ") + "      </Cells>\n";
        h += "      <PointData Scalars=\"temperature\" Vectors=\"velocity\">\n";
        for (const ArrayRef* a : std::initializer_list<const ArrayRef*>{&pres, &temp, &vel, &vmag, &r_domain, &r_pact,
                                                                         &r_vact, &r_tact})
            h += data_array_xml(*a, "        ");
        if (p->cfg.write_boundary_debug_arrays > 0) h += data_array_xml(r_nbc, "        ");
        h += "      </PointData>\n";
        h += "      <CellData Scalars=\"mat_id\">\n";
        for (const ArrayRef* a : std::initializer_list<const ArrayRef*>{&r_kind, &r_mat, &r_bc, &r_parent})
            h += data_array_xml(*a, "        ");
        h += "      </CellData>\n";
        h += "    </Piece>\n";
        h += "  </UnstructuredGrid>\n";
        h += "  <AppendedData encoding=\"raw\">\n   _";
        p->head = std::move(h);
        p->tail = "\n  </AppendedData>\n</VTKFile>\n";
    }

    // The dynamic arrays of one output with the reference's masks
    // (Post.cpp:975-1031): pressure on fluid-connected nodes, velocity and
    // its magnitude on fluid-only nodes, temperature everywhere.
    void encode_dynamic(cbs_post* p, const cbs_post_fields& f, std::vector<char>& out)
    {
        const int64_t N = p->N;
        out.resize(size_t(p->dynamic_bytes));
        char* base = out.data();
        const uint64_t n1 = uint64_t(N) * 8, n3 = uint64_t(N) * 24;
        double* pres = reinterpret_cast<double*>(base + 8);
        double* temp = reinterpret_cast<double*>(base + 16 + n1);
        double* vel = reinterpret_cast<double*>(base + 24 + 2 * n1);
        double* vmag = reinterpret_cast<double*>(base + 32 + 2 * n1 + n3);
        std::memcpy(base, &n1, 8);
        std::memcpy(base + 8 + n1, &n1, 8);
        std::memcpy(base + 16 + 2 * n1, &n3, 8);
        std::memcpy(base + 24 + 2 * n1 + n3, &n1, 8);
        const uint8_t* pa = p->pressure_active.data();
        const uint8_t* va = p->velocity_active.data();
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; ++i)
        {
            pres[i] = pa[i] ? safe_value(f.pres[i]) : 0.0;
            temp[i] = safe_value(f.temperature[i]);
            if (va[i])
            {
                vel[3 * i] = safe_value(f.unkno[3 * i]);
                vel[3 * i + 1] = safe_value(f.unkno[3 * i + 1]);
                vel[3 * i + 2] = safe_value(f.unkno[3 * i + 2]);
                vmag[i] = safe_value(f.velocity[i]);
            }
            else
            {
                vel[3 * i] = vel[3 * i + 1] = vel[3 * i + 2] = 0.0;
                vmag[i] = 0.0;
            }
        }
    }

    void write_vtu(cbs_post* p, int32_t iitime, const cbs_post_fields& f, std::vector<char>& dyn)
    {
        const auto t0 = std::chrono::steady_clock::now();
        fs::create_directories(p->dir);
        const fs::path final_path = p->dir / p->vtu_name(iitime);
        const fs::path tmp_path = fs::path(final_path.string() + ".part");
        encode_dynamic(p, f, dyn);
        const int fd = ::open(tmp_path.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
        if (fd < 0) throw PostFailure(CBS_ERR_IO, "Post::writeVTU - cannot open VTU file");
        try
        {
            write_all(fd, p->head.data(), p->head.size(), tmp_path.string());
            write_all(fd, dyn.data(), dyn.size(), tmp_path.string());
            write_all(fd, p->static_block.data(), p->static_block.size(), tmp_path.string());
            write_all(fd, p->tail.data(), p->tail.size(), tmp_path.string());
        }
        catch (...)
        {
            ::close(fd);
            throw;
        }
        if (::close(fd) != 0) throw PostFailure(CBS_ERR_IO, "Post::writeVTU - cannot write VTU file");
        fs::rename(tmp_path, final_path);
        const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::lock_guard<std::mutex> lk(p->mu);
        p->vtu_files += 1;
        p->vtu_bytes += int64_t(p->head.size() + dyn.size() + p->static_block.size() + p->tail.size());
        p->write_seconds += dt;
    }

    // Post::writePVD (Post.cpp:1133-1182) from an output history.
    void write_pvd(cbs_post* p, const std::vector<int32_t>& it, const std::vector<double>& t)
    {
        if (it.size() != t.size())
            throw PostFailure(CBS_ERR_BAD_ARGUMENT, "Post::writePVD - output history arrays are inconsistent");
        fs::create_directories(p->dir);
        std::string s;
        s += "<?xml version=\"1.0\"?>\n";
        s += "<VTKFile type=\"Collection\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
        s += "  <Collection>\n";
        for (size_t i = 0; i < it.size(); ++i)
        {
            s += "    <DataSet timestep=\"";
            put_g16(s, t[i]);
            s += "\" group=\"\" part=\"0\" file=\"" + p->vtu_name(it[i]) + "\"/>\n";
        }
        s += "  </Collection>\n";
        s += "</VTKFile>\n";
        const fs::path path = p->pvd_path();
        const fs::path tmp = fs::path(path.string() + ".part");
        FILE* f = std::fopen(tmp.c_str(), "wb");
        if (!f) throw PostFailure(CBS_ERR_IO, "Post::writePVD - cannot open PVD file");
        const bool ok = std::fwrite(s.data(), 1, s.size(), f) == s.size();
        if (std::fclose(f) != 0 || !ok) throw PostFailure(CBS_ERR_IO, "Post::writePVD - cannot open PVD file");
        fs::rename(tmp, path);
    }

    void worker_loop(cbs_post* p)
    {
        std::vector<char> dyn;
        for (;;)
        {
            Job job;
            {
                std::unique_lock<std::mutex> lk(p->mu);
                p->cv_work.wait(lk, [&] { return p->stop || !p->queue.empty(); });
                if (p->queue.empty()) return;  // stop requested and drained
                job = std::move(p->queue.front());
                p->queue.pop_front();
                p->busy = true;
            }
            int status = CBS_OK;
            std::string message;
            try
            {
                if (job.ready && job.ready(job.user) != 0)
                    throw PostFailure(CBS_ERR_HIP, "Post::writeVTU - the field snapshot failed");
                write_vtu(p, job.iitime, job.fields, dyn);
                write_pvd(p, job.hist_it, job.hist_t);
            }
            catch (const PostFailure& e)
            {
                status = e.status;
                message = e.what();
            }
            catch (const std::exception& e)
            {
                status = CBS_ERR_IO;
                message = e.what();
            }
            if (job.done) job.done(job.user);
            std::lock_guard<std::mutex> lk(p->mu);
            if (status != CBS_OK && p->bg_status == CBS_OK)
            {
                p->bg_status = status;
                p->bg_message = message;
            }
            p->busy = false;
            if (p->queue.empty()) p->cv_idle.notify_all();
        }
    }

    void open_csv(cbs_post* p, bool truncate, const char* who)
    {
        if (p->csv && !truncate) return;
        if (p->csv) std::fclose(p->csv);
        fs::create_directories(p->dir);
        p->csv = std::fopen(p->residual_path().c_str(), truncate ? "wb" : "ab");
        if (!p->csv) throw PostFailure(CBS_ERR_IO, std::string(who) + " - cannot open residual CSV file");
    }

    void write_solution_now(cbs_post* p, int32_t iitime, double rtime, const cbs_post_fields& f)
    {
        write_vtu(p, iitime, f, p->scratch);
        p->out_it.push_back(iitime);
        p->out_t.push_back(rtime);
        write_pvd(p, p->out_it, p->out_t);
    }

    void check_fields(const cbs_post* p, const cbs_post_fields* f)
    {
        if (!f || (p->N > 0 && (!f->pres || !f->temperature || !f->unkno || !f->velocity)))
            throw PostFailure(CBS_ERR_BAD_ARGUMENT, "cbs_post - missing solution array");
    }

    void drain(cbs_post* p)
    {
        std::unique_lock<std::mutex> lk(p->mu);
        p->cv_idle.wait(lk, [&] { return p->queue.empty() && !p->busy; });
    }
}  // namespace

extern "C" {

const char* cbs_post_last_error(void) { return g_last_error.c_str(); }

int cbs_post_create(const cbs_post_mesh* mesh, const cbs_post_config* cfg, const char* root, const char* case_name,
                    cbs_post** out)
{
    if (out) *out = nullptr;
    cbs_post* p = nullptr;
    const int rc = guarded([&] {
        if (!mesh || !cfg || !root || !case_name || !out || mesh->npoin < 0 || mesh->nelem < 0 || mesh->nboun < 0 ||
            (mesh->npoin > 0 && !mesh->coord) || (mesh->nelem > 0 && (!mesh->intma || !mesh->mat_elem)) ||
            (mesh->nboun > 0 && !mesh->iside) || !*case_name)
            throw PostFailure(CBS_ERR_BAD_ARGUMENT, "cbs_post_create - invalid argument");
        if (4 * int64_t(mesh->nelem) + 3 * int64_t(mesh->nboun) > INT32_MAX)
            throw PostFailure(CBS_ERR_BAD_ARGUMENT, "cbs_post_create - VTU offsets exceed Int32");
        p = new cbs_post;
        p->cfg = *cfg;
        p->N = mesh->npoin;
        p->E = mesh->nelem;
        p->B = mesh->nboun;
        p->dir = fs::path(root) / case_name;
        p->case_name = case_name;
        p->next_vtu_time = cfg->vtu_output_every_sim_time;
        build_static(p, *mesh);
        *out = p;
    });
    if (rc != CBS_OK) delete p;
    return rc;
}

int cbs_post_initialise_run_outputs(cbs_post* p, double rtime, const cbs_post_fields* f)
{
    if (!p) return CBS_ERR_BAD_ARGUMENT;
    return guarded([&] {
        drain(p);
        fs::create_directories(p->dir);
        p->out_it.clear();
        p->out_t.clear();
        p->next_vtu_time = p->cfg.vtu_output_every_sim_time;
        if (p->cfg.residual_log_enabled > 0)
        {
            open_csv(p, true, "Post::initialiseRunOutputs");
            static const char header[] =
                "iteration,time,dt,u_rel,u_norm,u_abs,v_rel,v_norm,v_abs,w_rel,w_norm,w_abs,p_rel,p_norm,p_abs,"
                "T_rel,T_norm,T_abs,velocity_rel_max,cg_iterations,cg_initial_l2,cg_final_l2,cg_relative_l2,"
                "cg_max_abs\n";
            if (std::fwrite(header, 1, sizeof header - 1, p->csv) != sizeof header - 1 || std::fflush(p->csv) != 0)
                throw PostFailure(CBS_ERR_IO, "Post::initialiseRunOutputs - cannot open residual CSV file");
        }
        if (p->cfg.vtu_output_enabled > 0)
        {
            check_fields(p, f);
            write_solution_now(p, 0, rtime, *f);
        }
    });
}

int cbs_post_write_residual_row(cbs_post* p, int32_t iitime, const cbs_scalars* sc)
{
    if (!p || !sc) return CBS_ERR_BAD_ARGUMENT;
    return guarded([&] {
        if (p->cfg.residual_log_enabled < 1) return;
        if (p->cfg.residual_log_every > 1 && (iitime % p->cfg.residual_log_every) != 0) return;
        open_csv(p, false, "Post::writeResidualRow");
        // Convergence::velocityResidual (Convergence.cpp:357-360): std::max
        // over {hb0, hb3, hb6}, which keeps the first of equal or unordered
        double vmax = sc->hb[0];
        if (vmax < sc->hb[3]) vmax = sc->hb[3];
        if (vmax < sc->hb[6]) vmax = sc->hb[6];
        std::string row = std::to_string(iitime);
        auto col = [&](double v) {
            row += ',';
            put_g16(row, safe_value(v));
        };
        col(sc->rtime);
        col(sc->dtreal);
        for (int k = 0; k < 15; ++k) col(sc->hb[k]);
        col(vmax);
        row += ',' + std::to_string(sc->last_cg_iterations);
        col(sc->last_cg_initial_l2);
        col(sc->last_cg_final_l2);
        col(sc->last_cg_relative_l2);
        col(sc->last_cg_max_abs);
        row += '\n';
        if (std::fwrite(row.data(), 1, row.size(), p->csv) != row.size() || std::fflush(p->csv) != 0)
            throw PostFailure(CBS_ERR_IO, "Post::writeResidualRow - cannot open residual CSV file");
    });
}

int cbs_post_should_write_vtu(cbs_post* p, int32_t iitime, double rtime, int32_t* due)
{
    if (!p || !due) return CBS_ERR_BAD_ARGUMENT;
    *due = 0;
    const cbs_post_config& c = p->cfg;
    if (c.vtu_output_enabled < 1) return CBS_OK;
    if (iitime == 0)
    {
        *due = 1;
        return CBS_OK;
    }
    // write_by_iteration || write_by_simulated_time (Post.cpp:563-604); the
    // second runs, and advances the next output time, only when the first
    // is false
    if (c.vtu_output_every_iterations > 0 && iitime % c.vtu_output_every_iterations == 0)
    {
        *due = 1;
        return CBS_OK;
    }
    if (c.transient_on < 1 || c.vtu_output_every_sim_time <= 0.0) return CBS_OK;
    if (rtime + 1.0e-14 < p->next_vtu_time) return CBS_OK;
    while (p->next_vtu_time <= rtime + 1.0e-14) p->next_vtu_time += c.vtu_output_every_sim_time;
    *due = 1;
    return CBS_OK;
}

int cbs_post_write_solution(cbs_post* p, int32_t iitime, double rtime, const cbs_post_fields* f)
{
    if (!p) return CBS_ERR_BAD_ARGUMENT;
    return guarded([&] {
        if (p->cfg.vtu_output_enabled < 1) return;
        check_fields(p, f);
        drain(p);  // keeps the PVD history in order
        write_solution_now(p, iitime, rtime, *f);
    });
}

int cbs_post_write_solution_async(cbs_post* p, int32_t iitime, double rtime, const cbs_post_fields* f,
                                  cbs_post_ready_fn ready, cbs_post_done_fn done, void* user)
{
    if (!p)
    {
        if (done) done(user);
        return CBS_ERR_BAD_ARGUMENT;
    }
    bool queued = false;
    const int rc = guarded([&] {
        if (p->cfg.vtu_output_enabled < 1) return;
        check_fields(p, f);
        Job job;
        job.iitime = iitime;
        job.fields = *f;
        job.ready = ready;
        job.done = done;
        job.user = user;
        p->out_it.push_back(iitime);
        p->out_t.push_back(rtime);
        job.hist_it = p->out_it;
        job.hist_t = p->out_t;
        std::lock_guard<std::mutex> lk(p->mu);
        if (!p->worker.joinable()) p->worker = std::thread(worker_loop, p);
        p->queue.push_back(std::move(job));
        queued = true;
        p->cv_work.notify_one();
    });
    if (!queued && done) done(user);
    return rc;
}

int cbs_post_last_output(cbs_post* p, int32_t* iitime)
{
    if (!p || !iitime) return CBS_ERR_BAD_ARGUMENT;
    *iitime = p->out_it.empty() ? -1 : p->out_it.back();
    return CBS_OK;
}

int cbs_post_flush(cbs_post* p)
{
    if (!p) return CBS_ERR_BAD_ARGUMENT;
    drain(p);
    std::lock_guard<std::mutex> lk(p->mu);
    const int s = p->bg_status;
    if (s != CBS_OK) g_last_error = p->bg_message;
    p->bg_status = CBS_OK;
    p->bg_message.clear();
    return s;
}

int cbs_post_stats(cbs_post* p, int64_t* vtu_files, int64_t* vtu_bytes, double* write_seconds)
{
    if (!p) return CBS_ERR_BAD_ARGUMENT;
    std::lock_guard<std::mutex> lk(p->mu);
    if (vtu_files) *vtu_files = p->vtu_files;
    if (vtu_bytes) *vtu_bytes = p->vtu_bytes;
    if (write_seconds) *write_seconds = p->write_seconds;
    return CBS_OK;
}

void cbs_post_destroy(cbs_post* p)
{
    if (!p) return;
    drain(p);
    {
        std::lock_guard<std::mutex> lk(p->mu);
        p->stop = true;
        p->cv_work.notify_all();
    }
    if (p->worker.joinable()) p->worker.join();
    if (p->csv) std::fclose(p->csv);
    delete p;
}

}  // extern "C"

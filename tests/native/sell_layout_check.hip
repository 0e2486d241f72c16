// Host check of the assembled-operator layout built by build_sell_host
// (cbs-3d_amd/csrc/hip/k_sell.hip): on a Kuhn-tet cube with solid elements
// and fixed nodes, element tiles of `chunk` shuffled fluid elements and an
// LDS budget of `lds` bytes, it checks the solver order (a permutation that
// keeps every segment's node range, rows of a tile consecutive), replays the
// value kernel and the apply kernel's indexing on the host with bounds
// checks, and compares the result with an element-by-element apply.  nseg =
// 2 builds two disjoint copies of the mesh as two segments.  Prints OK and
// exits 0 on success.
#include "../../cbs-3d_amd/csrc/hip/k_sell.hip"

#include <cmath>
#include <cstdio>
#include <random>

// device-side helpers the layout code does not use on the host
namespace cbsdev
{
    void hip_check(hipError_t, const char*) {}
    void launch_fill(cbs_hip_ctx*, double*, int64_t, double) {}
}
using namespace cbsdev;

int main(int argc, char** argv)
{
    const int n = argc > 1 ? atoi(argv[1]) : 7, chunk = argc > 2 ? atoi(argv[2]) : 150;
    const int lds = argc > 3 ? atoi(argv[3]) : 3200, nseg = argc > 4 ? atoi(argv[4]) : 1;
    const int nn = n + 1, N1 = nn * nn * nn, N = N1 * nseg;
    auto id = [&](int i, int j, int k) { return i + nn * (j + nn * k); };
    std::vector<int4> conn;
    std::vector<int> seg_elem{0};
    for (int sg = 0; sg < nseg; ++sg)
    {
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                {
                    int v[8];
                    for (int c = 0; c < 8; ++c) v[c] = sg * N1 + id(i + (c & 1), j + ((c >> 1) & 1), k + ((c >> 2) & 1));
                    const int perms[6][3] = {{1, 2, 4}, {1, 4, 2}, {2, 1, 4}, {2, 4, 1}, {4, 1, 2}, {4, 2, 1}};
                    for (auto& p : perms)
                    {
                        const int a = 0, b = a | p[0], c = b | p[1], d = c | p[2];
                        conn.push_back(make_int4(v[a], v[b], v[c], v[d]));
                    }
                }
        seg_elem.push_back(int(conn.size()));
    }
    const int E = int(conn.size());
    std::mt19937 rng(7);
    std::vector<int> mat(E, 0);
    for (int e = 0; e < E; ++e)
        if (rng() % 11 == 0) mat[e] = 1;
    std::vector<uint8_t> nflag(N, 0);
    for (int e = 0; e < E; ++e)
        if (!mat[e])
            for (int v : {conn[e].x, conn[e].y, conn[e].z, conn[e].w}) nflag[v] |= 1;
    for (int i = 0; i < N; ++i)
        if (rng() % 17 == 0) nflag[i] |= 4;
    // element tiles: chunks of shuffled fluid elements, per segment
    std::vector<int> fluid_all;
    std::vector<int32_t> elem_tslot(E, -1), tnp{0}, tn, segt{0};
    for (int sg = 0; sg < nseg; ++sg)
    {
        std::vector<int> fluid;
        for (int e = seg_elem[sg]; e < seg_elem[sg + 1]; ++e)
            if (!mat[e]) fluid.push_back(e);
        std::shuffle(fluid.begin(), fluid.end(), rng);
        for (size_t b = 0; b < fluid.size(); b += size_t(chunk))
        {
            std::vector<int> nodes;
            for (size_t k = b; k < std::min(fluid.size(), b + size_t(chunk)); ++k)
            {
                elem_tslot[fluid[k]] = int(fluid_all.size());
                fluid_all.push_back(fluid[k]);
                const int4 c = conn[fluid[k]];
                nodes.insert(nodes.end(), {c.x, c.y, c.z, c.w});
            }
            std::sort(nodes.begin(), nodes.end());
            nodes.erase(std::unique(nodes.begin(), nodes.end()), nodes.end());
            tn.insert(tn.end(), nodes.begin(), nodes.end());
            tnp.push_back(int(tn.size()));
        }
        segt.push_back(int(tnp.size()) - 1);
    }
    const int F = int(fluid_all.size()), T = int(tnp.size()) - 1;
    std::vector<int32_t> ip(N + 1, 0), inc(4 * size_t(E));
    for (int e = 0; e < E; ++e)
        for (int v : {conn[e].x, conn[e].y, conn[e].z, conn[e].w}) ip[v + 1]++;
    for (int i = 0; i < N; ++i) ip[i + 1] += ip[i];
    {
        std::vector<int> f(ip.begin(), ip.end() - 1);
        for (int e = 0; e < E; ++e)
        {
            const int v[4] = {conn[e].x, conn[e].y, conn[e].z, conn[e].w};
            for (int a = 0; a < 4; ++a) inc[f[v[a]]++] = e * 4 + a;
        }
    }
    SellSource s;
    s.N = N;
    s.F = F;
    s.ntiles = T;
    s.nseg = nseg;
    s.tile_node_ptr = tnp.data();
    s.tile_nodes = tn.data();
    s.seg_tile = segt.data();
    std::vector<int32_t> segn;
    for (int sg = 0; sg <= nseg; ++sg) segn.push_back(sg * N1);
    s.seg_node = segn.data();
    s.conn = conn.data();
    s.inc_ptr = ip.data();
    s.inc = inc.data();
    s.elem_tslot = elem_tslot.data();
    s.nflag = nflag.data();
    s.lds_bytes = lds;
    SellHost h;
    std::string err;
    if (!build_sell_host(s, h, err))
    {
        printf("build failed: %s\n", err.c_str());
        return 1;
    }
    const int RT = int(h.row_cnt.size());
    const int64_t entries = h.slice_off.back();
    printf("N=%d F=%d element tiles=%d row tiles=%d max staged=%d nnz=%lld entries=%lld contributions=%zu\n", N, F,
           T, RT, h.lmax, (long long)h.nnz, (long long)entries, h.cidx.size());
    int bad = 0;
    auto chk = [&](bool ok, const char* what) {
        if (!ok && bad++ < 10) printf("violation: %s\n", what);
    };
    chk(int(h.halo_ptr.size()) == RT + 1 && int(h.slice_ptr.size()) == RT + 1 && int(h.row_base.size()) == RT,
        "tile offset sizes");
    chk(int(h.gid.size()) == N, "solver order size");
    {
        std::vector<int> seen(N, 0);
        for (int i = 0; i < N; ++i)
        {
            const int g = h.gid[i];
            chk(g >= 0 && g < N && g / N1 == i / N1, "solver order keeps segment ranges");
            if (g >= 0 && g < N) seen[g]++;
        }
        for (int i = 0; i < N; ++i) chk(seen[i] == 1, "solver order is a permutation");
        for (int k = 0; k < nseg; ++k)  // a segment's tiles have consecutive rows from its first node on
        {
            int32_t next = k * N1;
            for (int t = h.seg_tile[k]; t < h.seg_tile[k + 1]; ++t)
            {
                chk(h.row_base[t] == next, "rows of consecutive tiles are consecutive");
                next += h.row_cnt[t];
            }
        }
        int64_t rows = 0;
        for (int t = 0; t < RT; ++t) rows += h.row_cnt[t];
        chk(rows == int64_t(h.rows.size()), "row count");
    }
    chk(int64_t(h.cols.size()) * 2 == entries && int64_t(h.cptr.size()) == entries + 1, "entry array sizes");
    chk(int(h.seg_tile.size()) == nseg + 1 && h.seg_tile[0] == 0 && h.seg_tile[nseg] == RT, "segment tile offsets");
    chk(h.cptr.back() == h.cidx.size(), "contribution offsets");
    chk(h.lmax <= 64 || int64_t(h.lmax) * 8 + int64_t(h.max_slices) * 8 <= lds, "LDS budget");
    for (uint32_t q : h.cidx) chk(int64_t(q) < 6LL * F, "contribution index");
    std::vector<double> g(size_t(6) * F);
    std::uniform_real_distribution<double> U(-1, 1);
    for (auto& v : g) v = U(rng);
    std::vector<double> v64(static_cast<size_t>(entries));
    for (int64_t q = 0; q < entries; ++q)
    {
        double a = 0;
        for (uint32_t k = h.cptr[q]; k < h.cptr[q + 1]; ++k) a += g[h.cidx[k]];
        v64[size_t(q)] = a;
    }
    std::vector<double> x(N), pdiag(N), y(N, -7.0), acc(N, 0.0);
    for (int i = 0; i < N; ++i)
    {
        x[i] = U(rng);
        pdiag[i] = 1 + U(rng);
    }
    // solver-order copies of x, pdiag and the flags, as the solve makes them
    std::vector<double> xs(N), ps(N), ys(N, -7.0);
    std::vector<uint8_t> fs(N);
    for (int i = 0; i < N; ++i)
    {
        xs[i] = x[h.gid[i]];
        ps[i] = pdiag[h.gid[i]];
        fs[i] = nflag[h.gid[i]];
    }
    std::vector<int> written(N, 0);
    double dot = 0;
    for (int t = 0; t < RT; ++t)
    {
        int sg = 0;
        while (t >= h.seg_tile[sg + 1]) ++sg;
        const int rb = h.row_base[t], nrows = h.row_cnt[t];
        chk(rb >= sg * N1 && rb + nrows <= (sg + 1) * N1, "rows outside the tile's segment");
        const int hb = h.halo_ptr[t], S = nrows + (h.halo_ptr[t + 1] - hb);
        chk(S <= h.lmax && S <= 65535, "staged count");
        std::vector<double> lx(S);
        for (int l = 0; l < S; ++l)
        {
            const int si = l < nrows ? rb + l : h.halo[hb + l - nrows];
            chk(si >= sg * N1 && si < (sg + 1) * N1, "staged node outside the tile's segment");
            lx[l] = xs[si];
        }
        const int s0 = h.slice_ptr[t], s1 = h.slice_ptr[t + 1];
        chk(s1 - s0 == (nrows + 63) / 64, "slices per tile");
        for (int sl = s0; sl < s1; ++sl)
        {
            const int64_t off = h.slice_off[sl];
            chk(off % 128 == 0, "slice alignment");
            const int np = int((h.slice_off[sl + 1] - off) >> 7);
            for (int lane = 0; lane < 64; ++lane)
            {
                double a = 0;
                for (int j = 0; j < np; ++j)
                {
                    const int64_t pi = (off >> 1) + int64_t(j) * 64 + lane;
                    const int64_t vi = off + 2 * (int64_t(j) * 64 + lane);
                    chk(pi < int64_t(h.cols.size()) && vi + 1 < entries, "entry index");
                    const uint32_t c = h.cols[size_t(pi)];
                    chk(int(c & 0xffff) < S && int(c >> 16) < S, "local column");
                    a += v64[size_t(vi)] * lx[c & 0xffff];
                    a += v64[size_t(vi) + 1] * lx[c >> 16];
                }
                const int row = (sl - s0) * 64 + lane;
                if (row < nrows)
                {
                    const int si = rb + row;
                    written[h.gid[si]]++;
                    double ap = 0;
                    if ((fs[si] & 1) && !(fs[si] & 4))
                    {
                        ap = a + ps[si] * lx[row];
                        dot += lx[row] * ap;
                    }
                    ys[si] = ap;
                }
            }
        }
    }
    for (int i = 0; i < N; ++i) y[h.gid[i]] = ys[i];
    // element-by-element reference (the tile-slot coefficient planes)
    const int plane[4][4] = {{-1, 0, 1, 2}, {0, -1, 3, 4}, {1, 3, -1, 5}, {2, 4, 5, -1}};
    for (int k = 0; k < F; ++k)
    {
        const int4 c = conn[fluid_all[k]];
        const int v[4] = {c.x, c.y, c.z, c.w};
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b)
                if (a != b) acc[v[a]] += g[size_t(plane[a][b]) * F + k] * x[v[b]];
    }
    double dref = 0, emax = 0, ymax = 0;
    int missing = 0, dup = 0;
    for (int i = 0; i < N; ++i)
    {
        if (!(nflag[i] & 1)) continue;
        if (written[i] == 0) missing++;
        if (written[i] > 1) dup++;
        const double yr = (nflag[i] & 4) ? 0.0 : acc[i] + pdiag[i] * x[i];
        dref += x[i] * yr;
        emax = std::max(emax, std::fabs(y[i] - yr));
        ymax = std::max(ymax, std::fabs(yr));
    }
    printf("max |y - y_element| = %.3e (max |y| %.3e), (x, y) %.15e vs %.15e, rows missing %d, rows written twice %d, "
           "violations %d\n",
           emax, ymax, dot, dref, missing, dup, bad);
    const bool ok = bad == 0 && missing == 0 && dup == 0 && emax <= 1e-12 * ymax &&
                    std::fabs(dot - dref) <= 1e-10 * std::fabs(dref);
    printf(ok ? "OK\n" : "FAIL\n");
    return ok ? 0 : 1;
}

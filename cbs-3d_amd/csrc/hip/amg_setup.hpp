// Host set-up of the smoothed-aggregation algebraic multigrid (SA-AMG)
// preconditioner of the pressure solve: plain C++ (no HIP), shared by the
// library and its host tests.
//
// Method (Vanek, Mandel and Brezina, "Algebraic multigrid by smoothed
// aggregation for second and fourth order elliptic problems", Computing 56,
// 1996):
//  - strength: j is a strong neighbour of i when |a_ij| >= eps_l
//    sqrt(a_ii a_jj), eps_0 = 0.08, halved on every coarser level;
//  - aggregation in three passes over the rows in order: (1) a row whose
//    strong neighbours are all unaggregated starts an aggregate with them;
//    (2) every row left joins the pass-1 aggregate of its strongest
//    aggregated strong neighbour; (3) rows still left start aggregates
//    with their unaggregated strong neighbours;
//  - tentative prolongator P0: the aggregate indicator (the constants, the
//    near-null space of a Laplacian);
//  - smoothed prolongator P = (I - w D^-1 A) P0, w = 4 / (3 lambda_max),
//    lambda_max of D^-1 A by power iteration (times 1.1);
//  - coarse operator A_c = P^T A P (Galerkin), by row-wise (Gustavson)
//    sparse products;
//  - levels until at most kCoarseMax unknowns remain; the coarsest is
//    inverted densely (Cholesky), so the V-cycle ends in an exact solve.
// Rows marked inactive (fixed or inactive pressure nodes) belong to no
// aggregate: their rows of P are empty, so the preconditioner is zero there.
// Everything is deterministic: passes run in row order, products sum in a
// fixed order.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <thread>
#include <vector>

namespace cbsamg
{
    struct Csr
    {
        int32_t n = 0, m = 0;  // rows, columns
        std::vector<int64_t> ptr{0};
        std::vector<int32_t> col;
        std::vector<double> val;
        int64_t nnz() const { return ptr.back(); }
    };

    struct Level
    {
        Csr A;                  // operator of this level (level 0: the fine operator)
        std::vector<double> dinv;
        double lambda = 0.0;    // lambda_max(D^-1 A) estimate
        double omega = 0.0;     // Jacobi weight of the smoother, 4 / (3 lambda)
        Csr P;                  // prolongation from the next level (n x n_next)
        Csr R;                  // restriction, P^T
    };

    struct Hierarchy
    {
        std::vector<Level> lv;       // lv.back() is the coarsest
        std::vector<double> inv;     // coarsest A^-1, dense n x n (symmetric)
    };

    constexpr int kCoarseMax = 1024;
    constexpr int kMaxLevels = 12;

    inline unsigned threads_for(int64_t work)
    {
        const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        return work < 100000 ? 1u : hw;
    }

    template <typename F>
    inline void parallel_rows(int32_t n, F f)
    {
        const unsigned nt = threads_for(n);
        if (nt == 1)
        {
            f(0, n);
            return;
        }
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; ++t)
            th.emplace_back([&, t] { f(int32_t(int64_t(n) * t / nt), int32_t(int64_t(n) * (t + 1) / nt)); });
        for (auto& x : th) x.join();
    }

    // y = A x
    inline void spmv(const Csr& A, const std::vector<double>& x, std::vector<double>& y)
    {
        y.assign(size_t(A.n), 0.0);
        parallel_rows(A.n, [&](int32_t b, int32_t e) {
            for (int32_t i = b; i < e; ++i)
            {
                double s = 0.0;
                for (int64_t k = A.ptr[i]; k < A.ptr[i + 1]; ++k) s += A.val[size_t(k)] * x[size_t(A.col[size_t(k)])];
                y[size_t(i)] = s;
            }
        });
    }

    inline std::vector<double> diagonal(const Csr& A)
    {
        std::vector<double> d(size_t(A.n), 0.0);
        for (int32_t i = 0; i < A.n; ++i)
            for (int64_t k = A.ptr[i]; k < A.ptr[i + 1]; ++k)
                if (A.col[size_t(k)] == i) d[size_t(i)] = A.val[size_t(k)];
        return d;
    }

    // lambda_max(D^-1 A) by 20 power iterations from a fixed start vector
    // (D-norm Rayleigh quotients), times 1.1.  Inactive rows (d = 0) are 0.
    inline double lambda_max(const Csr& A, const std::vector<double>& dinv)
    {
        std::vector<double> v(static_cast<size_t>(A.n)), w;
        uint64_t s = 0x9E3779B97F4A7C15ull;
        for (auto& x : v)
        {
            s = s * 6364136223846793005ull + 1442695040888963407ull;
            x = double((s >> 11) & 0xFFFFF) / double(0x100000) + 0.5;
        }
        for (size_t i = 0; i < v.size(); ++i)
            if (dinv[i] == 0.0) v[i] = 0.0;
        double lam = 0.0;
        for (int it = 0; it < 20; ++it)
        {
            spmv(A, v, w);
            double num = 0.0, den = 0.0, nrm = 0.0;
            for (size_t i = 0; i < v.size(); ++i)
            {
                num += v[i] * w[i];
                den += v[i] * v[i] / (dinv[i] == 0.0 ? 1.0 : dinv[i]);
                w[i] *= dinv[i];
                nrm += w[i] * w[i];
            }
            if (den > 0.0) lam = std::max(lam, num / den);
            if (nrm <= 0.0) break;
            const double sc = 1.0 / std::sqrt(nrm);
            for (size_t i = 0; i < v.size(); ++i) v[i] = w[i] * sc;
        }
        return 1.1 * lam;
    }

    // Aggregates of the active rows (agg[i] = -1 for inactive rows);
    // returns the number of aggregates.
    inline int32_t aggregate(const Csr& A, const std::vector<double>& d, const std::vector<uint8_t>& active,
                             double eps, std::vector<int32_t>& agg)
    {
        const int32_t n = A.n;
        agg.assign(size_t(n), -1);
        auto strong = [&](int32_t i, int64_t k) {
            const int32_t j = A.col[size_t(k)];
            if (j == i || !active[size_t(j)]) return false;
            return std::fabs(A.val[size_t(k)]) >= eps * std::sqrt(d[size_t(i)] * d[size_t(j)]);
        };
        constexpr int32_t kFree = -2;
        for (int32_t i = 0; i < n; ++i)
            if (active[size_t(i)]) agg[size_t(i)] = kFree;
        int32_t na = 0;
        // pass 1
        for (int32_t i = 0; i < n; ++i)
        {
            if (agg[size_t(i)] != kFree) continue;
            bool ok = true;
            for (int64_t k = A.ptr[i]; k < A.ptr[i + 1] && ok; ++k)
                if (strong(i, k) && agg[size_t(A.col[size_t(k)])] != kFree) ok = false;
            if (!ok) continue;
            agg[size_t(i)] = na;
            for (int64_t k = A.ptr[i]; k < A.ptr[i + 1]; ++k)
                if (strong(i, k)) agg[size_t(A.col[size_t(k)])] = na;
            ++na;
        }
        // pass 2: join the strongest pass-1 aggregate among strong neighbours
        std::vector<int32_t> first(agg);
        for (int32_t i = 0; i < n; ++i)
        {
            if (first[size_t(i)] != kFree) continue;
            double best = -1.0;
            int32_t pick = kFree;
            for (int64_t k = A.ptr[i]; k < A.ptr[i + 1]; ++k)
            {
                if (!strong(i, k)) continue;
                const int32_t a = first[size_t(A.col[size_t(k)])];
                if (a < 0) continue;
                const double w = std::fabs(A.val[size_t(k)]);
                if (w > best)
                {
                    best = w;
                    pick = a;
                }
            }
            if (pick >= 0) agg[size_t(i)] = pick;
        }
        // pass 3
        for (int32_t i = 0; i < n; ++i)
        {
            if (agg[size_t(i)] != kFree) continue;
            agg[size_t(i)] = na;
            for (int64_t k = A.ptr[i]; k < A.ptr[i + 1]; ++k)
                if (strong(i, k) && agg[size_t(A.col[size_t(k)])] == kFree) agg[size_t(A.col[size_t(k)])] = na;
            ++na;
        }
        return na;
    }

    // P = (I - w D^-1 A) P0 with P0 the aggregate indicator; columns of
    // each row ascending.
    inline Csr smoothed_prolongator(const Csr& A, const std::vector<double>& dinv, double w,
                                    const std::vector<int32_t>& agg, int32_t na)
    {
        const int32_t n = A.n;
        std::vector<std::vector<std::pair<int32_t, double>>> rows(static_cast<size_t>(n));
        parallel_rows(n, [&](int32_t b, int32_t e) {
            std::vector<std::pair<int32_t, double>> t;
            for (int32_t i = b; i < e; ++i)
            {
                if (agg[size_t(i)] < 0) continue;
                t.clear();
                t.push_back({agg[size_t(i)], 1.0});
                const double f = -w * dinv[size_t(i)];
                for (int64_t k = A.ptr[i]; k < A.ptr[i + 1]; ++k)
                {
                    const int32_t a = agg[size_t(A.col[size_t(k)])];
                    if (a >= 0) t.push_back({a, f * A.val[size_t(k)]});
                }
                std::stable_sort(t.begin(), t.end(),
                                 [](const auto& p, const auto& q) { return p.first < q.first; });
                auto& r = rows[size_t(i)];
                for (const auto& x : t)
                {
                    if (!r.empty() && r.back().first == x.first) r.back().second += x.second;
                    else r.push_back(x);
                }
            }
        });
        Csr P;
        P.n = n;
        P.m = na;
        for (int32_t i = 0; i < n; ++i)
        {
            for (const auto& x : rows[size_t(i)])
            {
                P.col.push_back(x.first);
                P.val.push_back(x.second);
            }
            P.ptr.push_back(int64_t(P.col.size()));
        }
        return P;
    }

    inline Csr transpose(const Csr& A)
    {
        Csr T;
        T.n = A.m;
        T.m = A.n;
        std::vector<int64_t> cnt(static_cast<size_t>(A.m) + 1, 0);
        for (int32_t c : A.col) cnt[size_t(c) + 1]++;
        for (int32_t j = 0; j < A.m; ++j) cnt[size_t(j) + 1] += cnt[size_t(j)];
        T.ptr = cnt;
        T.col.resize(A.col.size());
        T.val.resize(A.val.size());
        std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
        for (int32_t i = 0; i < A.n; ++i)  // rows ascending: columns of T ascending
            for (int64_t k = A.ptr[i]; k < A.ptr[i + 1]; ++k)
            {
                const int64_t p = pos[size_t(A.col[size_t(k)])]++;
                T.col[size_t(p)] = i;
                T.val[size_t(p)] = A.val[size_t(k)];
            }
        return T;
    }

    // C = A B (Gustavson, rows in parallel, columns of each row ascending).
    inline Csr multiply(const Csr& A, const Csr& B)
    {
        std::vector<std::vector<int32_t>> cc(static_cast<size_t>(A.n));
        std::vector<std::vector<double>> cv(static_cast<size_t>(A.n));
        parallel_rows(A.n, [&](int32_t b, int32_t e) {
            std::vector<double> acc(size_t(B.m), 0.0);
            std::vector<uint8_t> used(size_t(B.m), 0);
            std::vector<int32_t> list;
            for (int32_t i = b; i < e; ++i)
            {
                list.clear();
                for (int64_t k = A.ptr[i]; k < A.ptr[i + 1]; ++k)
                {
                    const int32_t j = A.col[size_t(k)];
                    const double a = A.val[size_t(k)];
                    for (int64_t q = B.ptr[j]; q < B.ptr[j + 1]; ++q)
                    {
                        const int32_t c = B.col[size_t(q)];
                        if (!used[size_t(c)])
                        {
                            used[size_t(c)] = 1;
                            list.push_back(c);
                        }
                        acc[size_t(c)] += a * B.val[size_t(q)];
                    }
                }
                std::sort(list.begin(), list.end());
                auto& rc = cc[size_t(i)];
                auto& rv = cv[size_t(i)];
                for (int32_t c : list)
                {
                    rc.push_back(c);
                    rv.push_back(acc[size_t(c)]);
                    acc[size_t(c)] = 0.0;
                    used[size_t(c)] = 0;
                }
            }
        });
        Csr C;
        C.n = A.n;
        C.m = B.m;
        for (int32_t i = 0; i < A.n; ++i)
        {
            C.col.insert(C.col.end(), cc[size_t(i)].begin(), cc[size_t(i)].end());
            C.val.insert(C.val.end(), cv[size_t(i)].begin(), cv[size_t(i)].end());
            C.ptr.push_back(int64_t(C.col.size()));
        }
        return C;
    }

    // Dense inverse of a symmetric positive definite matrix by Cholesky.
    inline bool dense_spd_inverse(const Csr& A, std::vector<double>& inv)
    {
        const int n = A.n;
        std::vector<double> L(static_cast<size_t>(n) * n, 0.0);
        for (int i = 0; i < n; ++i)
            for (int64_t k = A.ptr[i]; k < A.ptr[i + 1]; ++k) L[size_t(i) * n + size_t(A.col[size_t(k)])] = A.val[size_t(k)];
        for (int j = 0; j < n; ++j)
        {
            double dj = L[size_t(j) * n + j];
            for (int k = 0; k < j; ++k) dj -= L[size_t(j) * n + k] * L[size_t(j) * n + k];
            if (!(dj > 0.0)) return false;
            dj = std::sqrt(dj);
            L[size_t(j) * n + j] = dj;
            for (int i = j + 1; i < n; ++i)
            {
                double s = L[size_t(i) * n + j];
                for (int k = 0; k < j; ++k) s -= L[size_t(i) * n + k] * L[size_t(j) * n + k];
                L[size_t(i) * n + j] = s / dj;
            }
            for (int i = 0; i < j; ++i) L[size_t(i) * n + j] = 0.0;
        }
        // inverse column by column: L L^T x = e_c
        inv.assign(size_t(n) * n, 0.0);
        std::vector<double> y(static_cast<size_t>(n));
        for (int c = 0; c < n; ++c)
        {
            for (int i = 0; i < n; ++i)
            {
                double s = (i == c) ? 1.0 : 0.0;
                for (int k = 0; k < i; ++k) s -= L[size_t(i) * n + k] * y[size_t(k)];
                y[size_t(i)] = s / L[size_t(i) * n + i];
            }
            for (int i = n - 1; i >= 0; --i)
            {
                double s = y[size_t(i)];
                for (int k = i + 1; k < n; ++k) s -= L[size_t(k) * n + i] * inv[size_t(k) * n + c];
                inv[size_t(i) * n + c] = s / L[size_t(i) * n + i];
            }
        }
        return true;
    }

    // The hierarchy of A0 restricted to the active rows and columns (A0 must
    // hold no entry in an inactive row or column, and a positive diagonal on
    // every active row).
    inline bool setup(const Csr& A0, const std::vector<uint8_t>& active, Hierarchy& h, std::string& err)
    {
        h = Hierarchy{};
        h.lv.emplace_back();
        h.lv[0].A = A0;
        std::vector<uint8_t> act(active);
        double eps = 0.08;
        for (int l = 0;; ++l)
        {
            Level& L = h.lv[size_t(l)];
            const std::vector<double> d = diagonal(L.A);
            L.dinv.assign(d.size(), 0.0);
            for (size_t i = 0; i < d.size(); ++i)
                if (act[i])
                {
                    if (!(d[i] > 0.0))
                    {
                        err = "algebraic multigrid: non-positive diagonal at level " + std::to_string(l);
                        return false;
                    }
                    L.dinv[i] = 1.0 / d[i];
                }
            L.lambda = lambda_max(L.A, L.dinv);
            L.omega = L.lambda > 0.0 ? 4.0 / (3.0 * L.lambda) : 0.0;
            int64_t nact = 0;
            for (uint8_t a : act) nact += a ? 1 : 0;
            if (nact <= kCoarseMax || l + 1 >= kMaxLevels) break;
            std::vector<int32_t> agg;
            const int32_t na = aggregate(L.A, d, act, eps, agg);
            if (na == 0 || double(na) > 0.8 * double(nact)) break;  // coarsening stalled
            L.P = smoothed_prolongator(L.A, L.dinv, L.omega, agg, na);
            L.R = transpose(L.P);
            Csr AP = multiply(L.A, L.P);
            Level next;
            next.A = multiply(L.R, AP);
            h.lv.push_back(std::move(next));
            act.assign(size_t(na), 1);
            eps *= 0.5;
        }
        const Level& C = h.lv.back();
        // the coarsest level: dense inverse over its active rows (all rows
        // when it is a coarse level; level 0 alone keeps its inactive rows)
        if (C.A.n > 4 * kCoarseMax)
        {
            err = "algebraic multigrid: coarsening stalled at " + std::to_string(C.A.n) + " unknowns";
            return false;
        }
        Csr Cd = C.A;
        if (h.lv.size() == 1)
            for (int32_t i = 0; i < Cd.n; ++i)
                if (!act[size_t(i)])
                {
                    // identity on inactive rows; their entries in the inverse are zeroed below
                    Cd.col.insert(Cd.col.begin() + Cd.ptr[i], i);
                    Cd.val.insert(Cd.val.begin() + Cd.ptr[i], 1.0);
                    for (int32_t r = i + 1; r <= Cd.n; ++r) Cd.ptr[r] += 1;
                }
        if (!dense_spd_inverse(Cd, h.inv))
        {
            err = "algebraic multigrid: the coarsest operator is not positive definite";
            return false;
        }
        if (h.lv.size() == 1)
            for (int32_t i = 0; i < Cd.n; ++i)
                if (!act[size_t(i)])
                    for (int32_t j = 0; j < Cd.n; ++j)
                    {
                        h.inv[size_t(i) * Cd.n + j] = 0.0;
                        h.inv[size_t(j) * Cd.n + i] = 0.0;
                    }
        return true;
    }

    // One V-cycle z = M r on the host (the device's algorithm in FP64): one
    // damped-Jacobi sweep before and after the coarse correction on every
    // level, dense solve on the coarsest.  Used by the host tests.
    inline void vcycle(const Hierarchy& h, size_t l, const std::vector<double>& b, std::vector<double>& x)
    {
        const Level& L = h.lv[l];
        const size_t n = size_t(L.A.n);
        if (l + 1 == h.lv.size())
        {
            x.assign(n, 0.0);
            for (size_t i = 0; i < n; ++i)
            {
                double s = 0.0;
                for (size_t j = 0; j < n; ++j) s += h.inv[i * n + j] * b[j];
                x[i] = s;
            }
            return;
        }
        x.assign(n, 0.0);
        for (size_t i = 0; i < n; ++i) x[i] = L.omega * L.dinv[i] * b[i];
        std::vector<double> ax, r(n), bc, xc, px;
        spmv(L.A, x, ax);
        for (size_t i = 0; i < n; ++i) r[i] = b[i] - ax[i];
        spmv(L.R, r, bc);
        vcycle(h, l + 1, bc, xc);
        spmv(L.P, xc, px);
        for (size_t i = 0; i < n; ++i) x[i] += px[i];
        spmv(L.A, x, ax);
        for (size_t i = 0; i < n; ++i) x[i] += L.omega * L.dinv[i] * (b[i] - ax[i]);
    }
}

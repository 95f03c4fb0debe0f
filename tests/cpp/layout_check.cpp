// layout_check.cpp -- host-side checks of the HBM layouts the kernels read.
//
// Builds the plan of one system (host_setup.cpp, as ssa_ic_create does) and
//  * checks every index a kernel would form against its array (slices, lane
//    rows and positions, entry slots including the unguarded batch loads,
//    external columns, in-block references, walk steps with their prefetch
//    reach, walk records, LDS positions);
//  * runs the SpMV from its SELL-64 layout (row map included) and compares
//    it with the CSR product;
//  * emulates both sweeps lane by lane as the kernels do -- regular slices
//    with the shuffle resolution of in-block entries, the external part and
//    the step walk of oversized blocks read from the packed records (or the
//    level fallback) -- and compares them with a sequential substitution
//    over the host factor;
//  * combines the system with a 2-D Laplacian into one batch layout and
//    checks that the combined sweeps give each system's own results.
//
// usage: layout_check <matrix file> <complex 0|1> <block size> <colours> <use_abmc 0|1>
// matrix file: int32 n, int32 row_ptr[n+1], int32 col_idx[nnz], values[nnz]
// (double or complex<double>).  Prints "ok ..." and exits 0, or names the
// failed check and exits 1.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "host_plan.hpp"

using namespace sparsesolv_amd;

namespace {

[[noreturn]] void fail(const char* what, int line)
{
    std::printf("FAIL %s (line %d)\n", what, line);
    std::exit(1);
}
#define CHECK(c) \
    do {         \
        if (!(c)) fail(#c, __LINE__); \
    } while (0)

template <typename V>
std::vector<V> read_vec(FILE* f, size_t n)
{
    std::vector<V> v(n);
    if (n && std::fread(v.data(), sizeof(V), n, f) != n) throw std::runtime_error("short read");
    return v;
}

template <typename T>
double rel(const std::vector<T>& a, const std::vector<T>& b)
{
    double num = 0.0, den = 0.0;
    for (size_t i = 0; i < a.size(); ++i) {
        num += std::norm(a[i] - b[i]);
        den += std::norm(b[i]);
    }
    return den > 0.0 ? std::sqrt(num / den) : std::sqrt(num);
}

// one sweep from the layout, as the kernels compute it
template <typename T>
void emulate_sweep(const SweepLayout<T>& L, bool fwd, const std::vector<T>& invd,
                   const std::vector<T>& in, std::vector<T>& out, index_t n)
{
    const int BR = L.BR;   // 0: no regular slices (every block level scheduled)
    CHECK(BR >= 0 && BR <= 16 && (BR & (BR - 1)) == 0);
    CHECK(BR > 0 || L.nslices() == 0);
    for (index_t ph = 0; ph < L.nphases; ++ph) {
        // the distinct rows of earlier colours the phase's slices read
        std::set<index_t> ext;
        // regular slices
        for (index_t s = L.phase_slices[ph]; s < L.phase_slices[ph + 1]; ++s) {
            // the kernels find the slice's slots from the phase's first
            // slot and stride, its entry count K from the lane info words
            // (the same in every lane), and load slots of [o0, o0 + stride)
            // only: the first batch up to the stride, later ones up to K
            CHECK(size_t(ph) < L.phase_slot.size() && size_t(ph) < L.phase_stride.size());
            const index_t stride = L.phase_stride[ph];
            const index_t o0 = L.phase_slot[ph] + (s - L.phase_slices[ph]) * stride;
            CHECK(size_t(s) + 1 < L.slice_slot.size() && L.slice_slot[s] == o0);
            CHECK(size_t(s) * kWave + kWave <= L.lane_info.size());
            const index_t K = SweepLayout<T>::info_len(L.lane_info[size_t(s) * kWave]);
            for (int l = 0; l < kWave; ++l)
                CHECK(SweepLayout<T>::info_len(L.lane_info[size_t(s) * kWave + size_t(l)]) == K);
            const index_t o1 = o0 + K;
            CHECK(K >= 0 && K <= stride);
            CHECK(size_t(o0 + stride) * kWave <= L.val.size());
            CHECK(L.val.size() == L.col.size());
            for (index_t k = o1; k < o0 + stride; ++k)
                for (int l = 0; l < kWave; ++l) CHECK(L.col[size_t(k) * kWave + size_t(l)] == kPadCol);
            std::vector<T> acc(kWave), res(kWave, T(0)), xin(kWave), dinv(kWave);
            std::vector<std::vector<T>> intv(kWave, std::vector<T>(size_t(BR), T(0)));
            std::vector<index_t> rowv(kWave), posv(kWave);
            for (int l = 0; l < kWave; ++l) {
                const size_t li = size_t(s) * kWave + size_t(l);
                CHECK(li < L.lane_row.size() && li < L.lane_info.size());
                const index_t row = L.lane_row[li], pos = SweepLayout<T>::info_pos(L.lane_info[li]);
                rowv[l] = row;
                posv[l] = pos;
                CHECK(row < n);
                CHECK(pos >= 0 && pos < BR && l - pos >= 0);
                // the block starts in this lane's run of kBlockAlign lanes
                // (the split sweep resolves it within one wavefront)
                CHECK((l - pos) / kBlockAlign == l / kBlockAlign);
                const bool live = row >= 0;
                xin[l] = live ? in[size_t(row)] : T(0);
                dinv[l] = live ? invd[size_t(row)] : T(0);
                acc[l] = fwd ? xin[l] : T(0);
                for (index_t k = o0; k < o1; ++k) {
                    const size_t at = size_t(k) * kWave + size_t(l);
                    const index_t c = L.col[at];
                    if (c >= 0) {
                        CHECK(c < n);
                        ext.insert(c);
                        acc[l] -= L.val[at] * out[size_t(c)];
                    } else if (c != kPadCol) {
                        const index_t q = -1 - c;
                        CHECK(q >= 0 && q < BR - 1);
                        CHECK(!live || q < pos);
                        intv[size_t(l)][size_t(q)] = L.val[at];
                    }
                }
            }
            for (int j = 0; j < BR; ++j) {
                for (int l = 0; l < kWave; ++l)
                    if (rowv[l] >= 0 && posv[l] == j) {
                        res[l] = fwd ? acc[l] * dinv[l] : xin[l] + acc[l] * dinv[l];
                        out[size_t(rowv[l])] = res[l];
                    }
                if (j < BR - 1) {
                    std::vector<T> vj(kWave);
                    for (int l = 0; l < kWave; ++l) {
                        const int src = std::min(l - posv[l] + j, kWave - 1);
                        CHECK(src >= 0);
                        vj[l] = res[src];
                    }
                    for (int l = 0; l < kWave; ++l)
                        if (posv[l] > j) acc[l] -= intv[size_t(l)][size_t(j)] * vj[l];
                }
            }
        }
        CHECK(size_t(ph) < L.phase_ext_rows.size() && L.phase_ext_rows[ph] == int64_t(ext.size()));
        // oversized blocks of this phase
        const index_t b0 = L.phase_big[ph], b1 = L.phase_big[ph + 1];
        if (b1 == b0) continue;
        const int K = L.walk_kmax;
        if (K > 0) {
            const size_t FS = sizeof(T), W = size_t(L.walk_rec_words);
            CHECK(W == size_t(walk_record_words(K, int(FS))));
            const size_t nslots = L.walk_row.size();
            CHECK(nslots % kWave == 0 && L.walk_rec.size() == nslots / kWave * W * kWave);
            std::vector<T> sa(nslots, T(0)), sb(nslots, T(0));
            for (index_t bb = b0; bb < b1; ++bb) {
                const index_t s0 = L.walk_step_off[size_t(bb)], s1 = L.walk_step_off[size_t(bb) + 1];
                CHECK((s1 - s0) % kWalkDepth == 0);
                // the walk prefetches up to kWalkDepth steps past the block
                CHECK((size_t(s1) + kWalkDepth) * kWave <= nslots);
                // external part (the phase kernel)
                for (size_t t = size_t(s0) * kWave; t < size_t(s1) * kWave; ++t) {
                    const index_t row = L.walk_row[t];
                    if (row < 0) continue;
                    CHECK(row < n && t + 1 < L.walk_ext_rp.size());
                    T a = fwd ? in[size_t(row)] : T(0);
                    for (index_t k = L.walk_ext_rp[t]; k < L.walk_ext_rp[t + 1]; ++k) {
                        const index_t c = L.walk_ext_ci[size_t(k)];
                        CHECK(c >= 0 && c < n);
                        a -= L.walk_ext_v[size_t(k)] * out[size_t(c)];
                    }
                    sa[t] = a;
                    if (!fwd) sb[t] = in[size_t(row)];
                }
                // the step chain, operands decoded from the packed records
                std::vector<T> lds(size_t(L.max_big_rows) + 2, T(0));
                std::vector<char> done(size_t(L.max_big_rows) + 1, 0);
                done[size_t(L.max_big_rows)] = 1;   // the zero slot
                for (index_t st = s0; st < s1; ++st) {
                    std::vector<std::pair<index_t, T>> wr;
                    for (int l = 0; l < kWave; ++l) {
                        unsigned char rec[256];
                        for (size_t w = 0; w < W; ++w)
                            std::memcpy(rec + 16 * w, L.walk_rec[(size_t(st) * W + w) * kWave + size_t(l)].b, 16);
                        T v[8], d;
                        int ints[10];
                        std::memcpy(v, rec, size_t(K) * FS);
                        std::memcpy(&d, rec + size_t(K) * FS, FS);
                        std::memcpy(ints, rec + size_t(K + 1) * FS, size_t(2 + K) * 4);
                        const size_t t = size_t(st) * kWave + size_t(l);
                        CHECK(ints[0] == L.walk_row[t] && ints[1] == L.walk_pos[t]);
                        T acc = sa[t];
                        for (int k = 0; k < K; ++k) {
                            const int ix = ints[2 + k];
                            CHECK(ix >= 0 && ix <= L.max_big_rows);
                            CHECK(done[size_t(ix)] || v[k] == T(0));
                            acc -= v[k] * lds[size_t(ix)];
                        }
                        const T r = fwd ? acc * d : sb[t] + acc * d;
                        const int pos = ints[1], row = ints[0];
                        if (pos >= 0) {
                            CHECK(pos < L.max_big_rows && row >= 0 && row < n);
                            CHECK(d == invd[size_t(row)]);
                            wr.push_back({pos, r});
                            out[size_t(row)] = r;
                        }
                    }
                    for (auto& w : wr) {
                        lds[size_t(w.first)] = w.second;
                        done[size_t(w.first)] = 1;
                    }
                }
            }
        } else {
            // level fallback: rows grouped by dependency level
            for (index_t bb = b0; bb < b1; ++bb) {
                const index_t p0 = L.big_pos[size_t(bb)];
                const index_t l0 = L.big_lvl_off[size_t(bb)], l1 = L.big_lvl_off[size_t(bb) + 1];
                std::vector<T> lds;
                for (index_t lv = l0; lv + 1 < l1; ++lv)
                    for (index_t q = L.big_lvl[size_t(lv)]; q < L.big_lvl[size_t(lv) + 1]; ++q) {
                        const size_t gp = size_t(p0 + q);
                        const index_t row = L.big_rows[gp];
                        CHECK(row >= 0 && row < n);
                        T acc = fwd ? in[size_t(row)] : T(0);
                        for (index_t k = L.big_rp[gp]; k < L.big_rp[gp + 1]; ++k) {
                            const index_t c = L.big_ci[size_t(k)];
                            T xv;
                            if (c >= 0) {
                                CHECK(c < n);
                                xv = out[size_t(c)];
                            } else {
                                CHECK(size_t(-1 - c) < lds.size());
                                xv = lds[size_t(-1 - c)];
                            }
                            acc -= L.big_v[size_t(k)] * xv;
                        }
                        const T r = fwd ? acc * invd[size_t(row)] : in[size_t(row)] + acc * invd[size_t(row)];
                        out[size_t(row)] = r;
                        CHECK(size_t(q) == lds.size());
                        lds.push_back(r);
                    }
            }
        }
    }
}

template <typename T>
int run(const char* path, int block, int colours, bool abmc)
{
    FILE* f = std::fopen(path, "rb");
    if (!f) fail("open matrix file", __LINE__);
    int32_t n = 0;
    if (std::fread(&n, 4, 1, f) != 1) fail("read n", __LINE__);
    const auto rp = read_vec<int32_t>(f, size_t(n) + 1);
    const auto ci = read_vec<int32_t>(f, size_t(rp[size_t(n)]));
    const auto v = read_vec<T>(f, size_t(rp[size_t(n)]));
    std::fclose(f);
    IcSettings st;
    st.abmc_block_size = block;
    st.abmc_num_colors = colours;
    st.auto_shift = true;
    st.use_abmc = abmc;
    ssa_host::HostPlan<T, T> plan;
    ssa_host::build_plan<T, T>(plan, n, rp.data(), ci.data(), v.data(), st);
    const auto& ord = plan.fac.sched.ordering;   // device row of original row i
    CHECK(ord.size() == size_t(n));

    // inputs: a fixed pattern in device order
    std::vector<T> x(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) x[size_t(i)] = T(std::sin(0.37 * i + 0.11) + 0.5);

    // SpMV: row map a permutation, product equal to the CSR product
    const auto& M = plan.spmv;
    CHECK(M.row.size() == size_t(M.nslices) * kWave);
    std::vector<char> seen(size_t(n), 0);
    std::vector<T> y(static_cast<size_t>(n), T(0)), yref(static_cast<size_t>(n), T(0));
    for (index_t s = 0; s < M.nslices; ++s)
        for (int l = 0; l < kWave; ++l) {
            T acc = T(0);
            for (index_t k = M.slice_off[s]; k < M.slice_off[s + 1]; ++k) {
                const size_t at = size_t(k) * kWave + size_t(l);
                CHECK(at < M.val.size() && M.col[at] >= 0 && M.col[at] < n);
                acc += M.val[at] * x[size_t(M.col[at])];
            }
            const index_t row = M.row[size_t(s) * kWave + size_t(l)];
            if (row >= 0) {
                CHECK(row < n && !seen[size_t(row)]);
                seen[size_t(row)] = 1;
                y[size_t(row)] = acc;
            }
        }
    for (int i = 0; i < n; ++i) CHECK(seen[size_t(i)]);
    const auto& Ap = plan.fac.A_perm;
    for (int i = 0; i < n; ++i)
        for (int k = Ap.rp[size_t(i)]; k < Ap.rp[size_t(i) + 1]; ++k)
            yref[size_t(i)] += Ap.v[size_t(k)] * x[size_t(Ap.ci[size_t(k)])];
    const double e_spmv = rel(y, yref);
    CHECK(e_spmv < 1e-14);

    // sweeps from the layouts
    std::vector<T> t(static_cast<size_t>(n), T(0)), z(static_cast<size_t>(n), T(0));
    emulate_sweep(plan.fwd, true, plan.invd, x, t, n);
    emulate_sweep(plan.bwd, false, plan.invd, t, z, n);

    // sequential substitution over the host factor, in the factor's row
    // order (ABMC order; the original order on the level path)
    const auto& L = plan.fac.L;
    const auto& U = plan.fac.U;
    {
        // U is L^T, each row in ascending column order (a serial transpose)
        std::vector<index_t> rp(static_cast<size_t>(n) + 1, 0);
        for (index_t c : L.ci) rp[size_t(c) + 1]++;
        for (int i = 0; i < n; ++i) rp[size_t(i) + 1] += rp[size_t(i)];
        std::vector<index_t> ci(L.ci.size()), fill(rp.begin(), rp.end() - 1);
        std::vector<T> uv(L.v.size());
        for (int i = 0; i < n; ++i)
            for (int k = L.rp[size_t(i)]; k < L.rp[size_t(i) + 1]; ++k) {
                const index_t q = fill[size_t(L.ci[size_t(k)])]++;
                ci[size_t(q)] = i;
                uv[size_t(q)] = L.v[size_t(k)];
            }
        CHECK(U.rp == rp && U.ci == ci && U.v == uv);
    }
    const auto& d = plan.fac.inv_diag;
    const size_t N = static_cast<size_t>(n);
    std::vector<T> xf(N), tf(N), zf(N);
    auto dev = [&](int i) { return abmc ? i : ord[size_t(i)]; };
    for (int i = 0; i < n; ++i) xf[size_t(i)] = x[size_t(dev(i))];
    for (int i = 0; i < n; ++i) {
        T s = xf[size_t(i)];
        for (int k = L.rp[size_t(i)]; k < L.rp[size_t(i) + 1]; ++k)
            if (L.ci[size_t(k)] < i) s -= L.v[size_t(k)] * tf[size_t(L.ci[size_t(k)])];
        tf[size_t(i)] = s * d[size_t(i)];
    }
    for (int i = n - 1; i >= 0; --i) {
        T s = T(0);
        for (int k = U.rp[size_t(i)]; k < U.rp[size_t(i) + 1]; ++k)
            if (U.ci[size_t(k)] > i) s += U.v[size_t(k)] * zf[size_t(U.ci[size_t(k)])];
        zf[size_t(i)] = tf[size_t(i)] - d[size_t(i)] * s;
    }
    std::vector<T> tdev(N), zdev(N);
    for (int i = 0; i < n; ++i) {
        tdev[size_t(dev(i))] = tf[size_t(i)];
        zdev[size_t(dev(i))] = zf[size_t(i)];
    }
    const double e_fwd = rel(t, tdev), e_apply = rel(z, zdev);
    CHECK(e_fwd < 1e-12 && e_apply < 1e-12);

    // Batches: the combined layouts of this system, a 2-D Laplacian (other
    // colours, other strides) and this system again reproduce each
    // system's own sweeps bit for bit.
    const char* batch = "batch ok";
    {
        const int m = 23, n2 = m * m;
        std::vector<int32_t> rp2(1, 0), ci2;
        std::vector<T> v2;
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) {
                const int r = i * m + j;
                const int nb[5] = {r - m, r - 1, r, r + 1, r + m};
                const bool ok[5] = {i > 0, j > 0, true, j + 1 < m, i + 1 < m};
                for (int q = 0; q < 5; ++q)
                    if (ok[q]) {
                        ci2.push_back(nb[q]);
                        v2.push_back(T(q == 2 ? 4.0 : -1.0));
                    }
                rp2.push_back(int32_t(ci2.size()));
            }
        const int ns[3] = {n, n2, n};
        const int* rps[3] = {rp.data(), rp2.data(), rp.data()};
        const int* cis[3] = {ci.data(), ci2.data(), ci.data()};
        const T* vs[3] = {v.data(), v2.data(), v.data()};
        ssa_host::BatchPlan<T, T> bp;
        bool built = true;
        try {
            ssa_host::build_batch<T, T>(bp, 3, ns, rps, cis, vs, st);
        } catch (const std::runtime_error& e) {
            // oversized blocks past the walk, or wide blocks, are refused
            if (std::strncmp(e.what(), "batch:", 6) != 0) throw;
            built = false;
            batch = "batch refused";
        }
        if (built) {
            ssa_host::HostPlan<T, T> p2;
            ssa_host::build_plan<T, T>(p2, n2, rp2.data(), ci2.data(), v2.data(), st);
            std::vector<T> x2(static_cast<size_t>(n2)), t2(x2.size(), T(0)), z2(x2.size(), T(0));
            for (int i = 0; i < n2; ++i) x2[size_t(i)] = T(std::sin(0.37 * i + 0.11) + 0.5);
            emulate_sweep(p2.fwd, true, p2.invd, x2, t2, n2);
            emulate_sweep(p2.bwd, false, p2.invd, t2, z2, n2);
            const int NB = bp.row0.back();
            CHECK(bp.invd.size() == size_t(NB));
            std::vector<T> xb(size_t(NB), T(0)), tb(xb.size(), T(0)), zb(xb.size(), T(0));
            const std::vector<T>* xs[3] = {&x, &x2, &x};
            const std::vector<T>* ts[3] = {&t, &t2, &t};
            const std::vector<T>* zs[3] = {&z, &z2, &z};
            for (int q = 0; q < 3; ++q)
                for (int i = 0; i < ns[q]; ++i) xb[size_t(bp.row0[q] + i)] = (*xs[q])[size_t(i)];
            emulate_sweep(bp.fwd, true, bp.invd, xb, tb, NB);
            emulate_sweep(bp.bwd, false, bp.invd, tb, zb, NB);
            for (int q = 0; q < 3; ++q)
                for (int i = 0; i < ns[q]; ++i) {
                    CHECK(tb[size_t(bp.row0[q] + i)] == (*ts[q])[size_t(i)]);
                    CHECK(zb[size_t(bp.row0[q] + i)] == (*zs[q])[size_t(i)]);
                }
        }
    }
    std::printf("ok n=%d phases=%d BR=%d big=%d big_rows=%d levels=%d walk_k=%d/%d "
                "spmv %.1e forward %.1e apply %.1e %s\n",
                n, int(plan.fwd.nphases), plan.fwd.BR, int(plan.fwd.phase_big.back()),
                int(plan.fwd.max_big_rows), int(plan.fwd.max_levels), plan.fwd.walk_kmax,
                plan.bwd.walk_kmax, e_spmv, e_fwd, e_apply, batch);
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 6) {
        std::fprintf(stderr, "usage: %s matrix complex block colours use_abmc\n", argv[0]);
        return 2;
    }
    const bool cplx = std::atoi(argv[2]) != 0;
    const int block = std::atoi(argv[3]), colours = std::atoi(argv[4]);
    const bool abmc = std::atoi(argv[5]) != 0;
    try {
        return cplx ? run<std::complex<double>>(argv[1], block, colours, abmc)
                    : run<double>(argv[1], block, colours, abmc);
    } catch (const std::exception& e) {
        std::printf("FAIL %s\n", e.what());
        return 1;
    }
}

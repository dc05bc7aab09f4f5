// Host-side data plumbing: rating-matrix container, reference .bin I/O,
// CSR -> CSC transpose and the seeded synthetic workload generator.
#include "als_data.h"
#include "als_status.h"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <unordered_set>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

struct als_dataset {
    int64_t m = 0, n = 0;
    std::vector<int> csr_ptr, csr_idx, csc_ptr, csc_idx, coo_row, t_row, t_col;
    std::vector<float> csr_val, csc_val, t_val;
};

static thread_local std::string g_err;
const char *als_data_last_error(void) { return g_err.c_str(); }
static int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

// ---------------------------------------------------------------- RNG
// splitmix64 seeds independent xoshiro256** streams per (purpose, index), so
// every row draws from its own stream and the output does not depend on how
// rows are spread over threads.
static inline uint64_t splitmix64(uint64_t &s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
struct Rng {
    uint64_t s[4];
    bool has_spare = false;
    double spare = 0.0;
    Rng(uint64_t seed, uint64_t purpose, uint64_t index) {
        uint64_t x = seed ^ (purpose * 0xd1b54a32d192ed03ULL) ^ (index * 0x8cb92ba72f3d8dd7ULL);
        for (auto &v : s) v = splitmix64(x);
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    double uniform_pos() { return ((double)(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }  // (0,1)
    uint64_t below(uint64_t n) { return (uint64_t)(uniform() * (double)n) % n; }
    double normal() {  // Box-Muller
        if (has_spare) { has_spare = false; return spare; }
        double u1 = uniform_pos(), u2 = uniform();
        double r = std::sqrt(-2.0 * std::log(u1)), th = 6.283185307179586 * u2;
        spare = r * std::sin(th);
        has_spare = true;
        return r * std::cos(th);
    }
};
enum : uint64_t { P_ROWSIZE = 1, P_ROWITEMS = 2, P_UFAC = 3, P_VFAC = 4, P_SPLIT = 5, P_PAIRS = 6, P_ADJUST = 7 };

// ---------------------------------------------------------------- transpose
int als_csr_to_csc(int64_t rows, int64_t cols, int64_t nnz, const int *ptr,
                   const int *idx, const float *val, int *cptr, int *cidx, float *cval) {
    if (rows < 0 || cols < 0 || nnz < 0) return fail(ALS_E_INVALID, "negative size");
    if ((int64_t)ptr[rows] != nnz) return fail(ALS_E_INVALID, "row_ptr[rows] != nnz");
    std::vector<int64_t> cnt(cols + 1, 0);
    for (int64_t k = 0; k < nnz; k++) {
        int c = idx[k];
        if (c < 0 || c >= cols) return fail(ALS_E_INVALID, "column index out of range");
        cnt[c + 1]++;
    }
    for (int64_t c = 0; c < cols; c++) cnt[c + 1] += cnt[c];
    for (int64_t c = 0; c <= cols; c++) cptr[c] = (int)cnt[c];
    for (int64_t r = 0; r < rows; r++)
        for (int64_t k = ptr[r]; k < ptr[r + 1]; k++) {
            int64_t dst = cnt[idx[k]]++;
            cidx[dst] = (int)r;
            cval[dst] = val[k];
        }
    return ALS_OK;
}

static int finish_dataset(als_dataset *ds) {
    int64_t nnz = (int64_t)ds->csr_idx.size();
    ds->csc_ptr.assign(ds->n + 1, 0);
    ds->csc_idx.resize(nnz);
    ds->csc_val.resize(nnz);
    int rc = als_csr_to_csc(ds->m, ds->n, nnz, ds->csr_ptr.data(), ds->csr_idx.data(),
                            ds->csr_val.data(), ds->csc_ptr.data(), ds->csc_idx.data(),
                            ds->csc_val.data());
    if (rc) return rc;
    ds->coo_row.resize(nnz);
    for (int64_t r = 0; r < ds->m; r++)
        for (int64_t k = ds->csr_ptr[r]; k < ds->csr_ptr[r + 1]; k++) ds->coo_row[k] = (int)r;
    return ALS_OK;
}

// ---------------------------------------------------------------- generator
static double quantise(double x, const als_synth_spec &s) {
    double q = std::nearbyint(x / s.rating_step) * s.rating_step;
    return std::min(s.rating_max, std::max(s.rating_min, q));
}

int als_synth_generate(const als_synth_spec *sp, als_dataset **out) {
    if (!sp || !out) return fail(ALS_E_INVALID, "null argument");
    const als_synth_spec s = *sp;
    if (s.m <= 0 || s.n <= 0 || s.total <= 0 || s.n_test < 0 || s.n_test > s.total)
        return fail(ALS_E_INVALID, "bad sizes");
    if (s.total > (int64_t)s.m * s.n) return fail(ALS_E_INVALID, "more ratings than cells");
    if (s.total >= (int64_t)INT32_MAX) return fail(ALS_E_INVALID, "nnz must fit int32 row pointers");
    if (s.rank <= 0 || s.rating_step <= 0) return fail(ALS_E_INVALID, "bad rank or step");

    const int64_t m = s.m, n = s.n;
    // Item popularity CDF, Zipf(s) over item index.
    std::vector<double> cdf(n);
    double acc = 0.0;
    for (int64_t i = 0; i < n; i++) { acc += std::pow((double)(i + 1), -s.zipf_s); cdf[i] = acc; }
    for (auto &c : cdf) c /= acc;
    auto zipf = [&](Rng &r) -> int64_t {
        double u = r.uniform();
        int64_t i = std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin();
        return std::min<int64_t>(i, n - 1);
    };

    // Ground-truth factors, coordinates ~ N(0, rank^-1/2) so Var(u.v) = 1.
    const int K = s.rank;
    const double sd = std::pow((double)K, -0.25);
    std::vector<double> Ut((size_t)m * K), Vt((size_t)n * K);
#pragma omp parallel for schedule(static)
    for (int64_t u = 0; u < m; u++) { Rng r(s.seed, P_UFAC, u); for (int k = 0; k < K; k++) Ut[u * K + k] = sd * r.normal(); }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) { Rng r(s.seed, P_VFAC, i); for (int k = 0; k < K; k++) Vt[i * K + k] = sd * r.normal(); }

    // All ratings, grouped by row, items ascending inside a row.
    std::vector<int64_t> rptr(m + 1, 0);
    std::vector<int> items;
    std::vector<float> vals;

    if (s.row_model == ALS_SYNTH_UNIFORM_PAIRS) {
        Rng r(s.seed, P_PAIRS, 0);
        std::unordered_set<int64_t> seen;
        seen.reserve((size_t)s.total * 2);
        std::vector<std::pair<int64_t, int>> pairs;
        pairs.reserve(s.total);
        while ((int64_t)pairs.size() < s.total) {
            int64_t u = (int64_t)r.below(m), i = zipf(r);
            if (seen.insert(u * n + i).second) pairs.push_back({u, (int)i});
        }
        std::vector<double> noise(s.total);
        for (auto &e : noise) e = r.normal();
        std::vector<int64_t> order(s.total);
        for (int64_t t = 0; t < s.total; t++) order[t] = t;
        std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
            return pairs[a].first != pairs[b].first ? pairs[a].first < pairs[b].first
                                                    : pairs[a].second < pairs[b].second;
        });
        items.resize(s.total);
        vals.resize(s.total);
        for (int64_t t = 0; t < s.total; t++) {
            int64_t o = order[t], u = pairs[o].first;
            int i = pairs[o].second;
            rptr[u + 1]++;
            items[t] = i;
            double d = 0.0;
            for (int k = 0; k < K; k++) d += Ut[u * K + k] * Vt[(int64_t)i * K + k];
            vals[t] = (float)quantise(d + s.rating_offset + s.noise_sigma * noise[o], s);
        }
        for (int64_t u = 0; u < m; u++) rptr[u + 1] += rptr[u];
    } else if (s.row_model == ALS_SYNTH_LOGNORMAL_ROWS) {
        const int64_t rmin = std::max<int64_t>(0, s.row_min), rmax = std::min<int64_t>(s.row_max, n);
        if (rmin > rmax) return fail(ALS_E_INVALID, "row_min > row_max");
        if (rmin * m > s.total || rmax * m < s.total) return fail(ALS_E_INVALID, "total out of reach of the row bounds");
        std::vector<int64_t> cnt(m);
#pragma omp parallel for schedule(static)
        for (int64_t u = 0; u < m; u++) {
            Rng r(s.seed, P_ROWSIZE, u);
            double x = std::exp(s.ln_mu + s.ln_sigma * r.normal());
            cnt[u] = std::min(rmax, std::max(rmin, (int64_t)std::llround(x)));
        }
        // Scale to the exact total, then settle the remainder one rating at a
        // time over a seeded permutation of the rows.
        int64_t S = 0;
        for (auto c : cnt) S += c;
        if (S != s.total) {
            double scale = (double)s.total / (double)S;
            S = 0;
            for (auto &c : cnt) { c = std::min(rmax, std::max(rmin, (int64_t)std::llround(c * scale))); S += c; }
        }
        std::vector<int64_t> perm(m);
        for (int64_t u = 0; u < m; u++) perm[u] = u;
        { Rng r(s.seed, P_ADJUST, 0); for (int64_t u = m - 1; u > 0; u--) std::swap(perm[u], perm[r.below(u + 1)]); }
        while (S != s.total) {
            int64_t step = S < s.total ? 1 : -1;
            bool moved = false;
            for (int64_t t = 0; t < m && S != s.total; t++) {
                int64_t &c = cnt[perm[t]];
                if (c + step < rmin || c + step > rmax) continue;
                c += step; S += step; moved = true;
            }
            if (!moved) return fail(ALS_E_INVALID, "cannot reach the total");
        }
        for (int64_t u = 0; u < m; u++) rptr[u + 1] = rptr[u] + cnt[u];
        items.resize(s.total);
        vals.resize(s.total);
        // Items per row: without replacement, probability proportional to the
        // Zipf weight. Sparse rows: rejection of repeats; dense rows:
        // exponential keys -log(u)/w (Efraimidis-Spirakis), k smallest.
#pragma omp parallel
        {
            std::vector<char> mark(n, 0);
            std::vector<std::pair<double, int>> keys;
#pragma omp for schedule(dynamic, 256)
            for (int64_t u = 0; u < m; u++) {
                Rng r(s.seed, P_ROWITEMS, u);
                int64_t k = cnt[u], base = rptr[u];
                int *it = items.data() + base;
                if (k == n) {
                    for (int64_t i = 0; i < n; i++) it[i] = (int)i;
                } else if (k * 4 > n) {
                    keys.resize(n);
                    for (int64_t i = 0; i < n; i++)
                        keys[i] = {-std::log(r.uniform_pos()) * std::pow((double)(i + 1), s.zipf_s), (int)i};
                    std::nth_element(keys.begin(), keys.begin() + k, keys.end());
                    for (int64_t t = 0; t < k; t++) it[t] = keys[t].second;
                } else {
                    int64_t got = 0;
                    while (got < k) {
                        int64_t i = zipf(r);
                        if (!mark[i]) { mark[i] = 1; it[got++] = (int)i; }
                    }
                    for (int64_t t = 0; t < k; t++) mark[it[t]] = 0;
                }
                std::sort(it, it + k);
                for (int64_t t = 0; t < k; t++) {
                    int64_t i = it[t];
                    double d = 0.0;
                    for (int q = 0; q < K; q++) d += Ut[u * K + q] * Vt[i * K + q];
                    vals[base + t] = (float)quantise(d + s.rating_offset + s.noise_sigma * r.normal(), s);
                }
            }
        }
    } else {
        return fail(ALS_E_INVALID, "unknown row model");
    }

    // Held-out split: exactly n_test ratings, uniformly (Floyd's algorithm).
    std::vector<char> is_test(s.total, 0);
    {
        Rng r(s.seed, P_SPLIT, 0);
        for (int64_t j = s.total - s.n_test; j < s.total; j++) {
            int64_t t = (int64_t)r.below(j + 1);
            if (is_test[t]) is_test[j] = 1; else is_test[t] = 1;
        }
    }
    auto *ds = new als_dataset;
    ds->m = m;
    ds->n = n;
    int64_t ntr = s.total - s.n_test;
    ds->csr_ptr.assign(m + 1, 0);
    ds->csr_idx.reserve(ntr);
    ds->csr_val.reserve(ntr);
    ds->t_row.reserve(s.n_test);
    ds->t_col.reserve(s.n_test);
    ds->t_val.reserve(s.n_test);
    for (int64_t u = 0; u < m; u++) {
        for (int64_t t = rptr[u]; t < rptr[u + 1]; t++) {
            if (is_test[t]) {
                ds->t_row.push_back((int)u); ds->t_col.push_back(items[t]); ds->t_val.push_back(vals[t]);
            } else {
                ds->csr_idx.push_back(items[t]); ds->csr_val.push_back(vals[t]);
            }
        }
        ds->csr_ptr[u + 1] = (int)ds->csr_idx.size();
    }
    int rc = finish_dataset(ds);
    if (rc) { delete ds; return rc; }
    *out = ds;
    return ALS_OK;
}

int als_dataset_from_csr(int64_t m, int64_t n, int64_t nnz, const int *ptr,
                         const int *idx, const float *val, int64_t n_test,
                         const int *t_row, const int *t_col, const float *t_val,
                         als_dataset **out) {
    if (!out || m < 0 || n < 0 || nnz < 0 || n_test < 0) return fail(ALS_E_INVALID, "bad arguments");
    if (ptr[0] != 0 || ptr[m] != nnz) return fail(ALS_E_INVALID, "row_ptr must start at 0 and end at nnz");
    for (int64_t r = 0; r < m; r++)
        if (ptr[r + 1] < ptr[r]) return fail(ALS_E_INVALID, "row_ptr must be non-decreasing");
    for (int64_t t = 0; t < n_test; t++)
        if (t_row[t] < 0 || t_row[t] >= m || t_col[t] < 0 || t_col[t] >= n)
            return fail(ALS_E_INVALID, "test index out of range");
    auto *ds = new als_dataset;
    ds->m = m;
    ds->n = n;
    ds->csr_ptr.assign(ptr, ptr + m + 1);
    ds->csr_idx.assign(idx, idx + nnz);
    ds->csr_val.assign(val, val + nnz);
    ds->t_row.assign(t_row, t_row + n_test);
    ds->t_col.assign(t_col, t_col + n_test);
    ds->t_val.assign(t_val, t_val + n_test);
    int rc = finish_dataset(ds);
    if (rc) { delete ds; return rc; }
    *out = ds;
    return ALS_OK;
}

// ---------------------------------------------------------------- .bin I/O
// Reference: host_sparse_matrix::malloc_load_file (cuda_sparse_matrix.cpp:6-27)
// rejects missing, empty and wrong-size files; the same checks apply here.
template <class T>
static int load_file(const std::string &path, std::vector<T> &dst, int64_t count) {
    std::ifstream ifs(path, std::ios::binary | std::ios::ate);
    if (!ifs) return fail(ALS_E_IO, path + ": " + std::strerror(errno));
    int64_t size = (int64_t)ifs.tellg();
    if (size == 0 && count > 0) return fail(ALS_E_IO, path + " is empty");
    if (size != count * (int64_t)sizeof(T)) return fail(ALS_E_IO, path + ": file size does not match buffer size");
    ifs.seekg(0, std::ios::beg);
    dst.resize(count);
    if (count && !ifs.read(reinterpret_cast<char *>(dst.data()), size)) return fail(ALS_E_IO, path + ": read failed");
    return ALS_OK;
}
template <class T>
static int save_file(const std::string &path, const std::vector<T> &src) {
    std::ofstream ofs(path, std::ios::binary);
    if (!ofs) return fail(ALS_E_IO, path + ": " + std::strerror(errno));
    if (!src.empty()) ofs.write(reinterpret_cast<const char *>(src.data()), src.size() * sizeof(T));
    return ofs ? ALS_OK : fail(ALS_E_IO, path + ": write failed");
}

int als_dataset_load_bin(const char *folder, int64_t m, int64_t n, int64_t nnz_train,
                         int64_t nnz_test, als_dataset **out) {
    if (!folder || !out) return fail(ALS_E_INVALID, "null argument");
    std::string d(folder);
    auto *ds = new als_dataset;
    ds->m = m;
    ds->n = n;
    int rc = 0;
    if (!rc) rc = load_file(d + "/R_train_csr.data.bin", ds->csr_val, nnz_train);
    if (!rc) rc = load_file(d + "/R_train_csr.indptr.bin", ds->csr_ptr, m + 1);
    if (!rc) rc = load_file(d + "/R_train_csr.indices.bin", ds->csr_idx, nnz_train);
    if (!rc) rc = load_file(d + "/R_train_csc.data.bin", ds->csc_val, nnz_train);
    if (!rc) rc = load_file(d + "/R_train_csc.indices.bin", ds->csc_idx, nnz_train);
    if (!rc) rc = load_file(d + "/R_train_csc.indptr.bin", ds->csc_ptr, n + 1);
    if (!rc) rc = load_file(d + "/R_test_coo.data.bin", ds->t_val, nnz_test);
    if (!rc) rc = load_file(d + "/R_test_coo.row.bin", ds->t_row, nnz_test);
    if (!rc) rc = load_file(d + "/R_test_coo.col.bin", ds->t_col, nnz_test);
    if (!rc && (ds->csr_ptr[0] != 0 || ds->csr_ptr[m] != nnz_train || ds->csc_ptr[0] != 0 || ds->csc_ptr[n] != nnz_train))
        rc = fail(ALS_E_INVALID, "index pointers do not span nnz_train");
    if (!rc) {
        for (int64_t k = 0; k < nnz_train && !rc; k++)
            if (ds->csr_idx[k] < 0 || ds->csr_idx[k] >= n || ds->csc_idx[k] < 0 || ds->csc_idx[k] >= m)
                rc = fail(ALS_E_INVALID, "train index out of range");
        for (int64_t t = 0; t < nnz_test && !rc; t++)
            if (ds->t_row[t] < 0 || ds->t_row[t] >= m || ds->t_col[t] < 0 || ds->t_col[t] >= n)
                rc = fail(ALS_E_INVALID, "test index out of range");
    }
    if (rc) { delete ds; return rc; }
    ds->coo_row.resize(nnz_train);
    for (int64_t r = 0; r < m; r++)
        for (int64_t k = ds->csr_ptr[r]; k < ds->csr_ptr[r + 1]; k++) ds->coo_row[k] = (int)r;
    *out = ds;
    return ALS_OK;
}

int als_dataset_save_bin(const als_dataset *ds, const char *folder) {
    if (!ds || !folder) return fail(ALS_E_INVALID, "null argument");
    std::string d(folder);
    int rc = 0;
    if (!rc) rc = save_file(d + "/R_train_csr.data.bin", ds->csr_val);
    if (!rc) rc = save_file(d + "/R_train_csr.indptr.bin", ds->csr_ptr);
    if (!rc) rc = save_file(d + "/R_train_csr.indices.bin", ds->csr_idx);
    if (!rc) rc = save_file(d + "/R_train_csc.data.bin", ds->csc_val);
    if (!rc) rc = save_file(d + "/R_train_csc.indices.bin", ds->csc_idx);
    if (!rc) rc = save_file(d + "/R_train_csc.indptr.bin", ds->csc_ptr);
    if (!rc) rc = save_file(d + "/R_train_coo.row.bin", ds->coo_row);
    if (!rc) rc = save_file(d + "/R_test_coo.data.bin", ds->t_val);
    if (!rc) rc = save_file(d + "/R_test_coo.row.bin", ds->t_row);
    if (!rc) rc = save_file(d + "/R_test_coo.col.bin", ds->t_col);
    return rc;
}

int als_dataset_info(const als_dataset *ds, int64_t *m, int64_t *n, int64_t *nnz_train, int64_t *nnz_test) {
    if (!ds) return fail(ALS_E_INVALID, "null dataset");
    if (m) *m = ds->m;
    if (n) *n = ds->n;
    if (nnz_train) *nnz_train = (int64_t)ds->csr_idx.size();
    if (nnz_test) *nnz_test = (int64_t)ds->t_row.size();
    return ALS_OK;
}
int als_dataset_csr(const als_dataset *ds, const int **p, const int **i, const float **v) {
    if (!ds) return fail(ALS_E_INVALID, "null dataset");
    *p = ds->csr_ptr.data(); *i = ds->csr_idx.data(); *v = ds->csr_val.data();
    return ALS_OK;
}
int als_dataset_csc(const als_dataset *ds, const int **p, const int **i, const float **v) {
    if (!ds) return fail(ALS_E_INVALID, "null dataset");
    *p = ds->csc_ptr.data(); *i = ds->csc_idx.data(); *v = ds->csc_val.data();
    return ALS_OK;
}
int als_dataset_test(const als_dataset *ds, const int **r, const int **c, const float **v) {
    if (!ds) return fail(ALS_E_INVALID, "null dataset");
    *r = ds->t_row.data(); *c = ds->t_col.data(); *v = ds->t_val.data();
    return ALS_OK;
}
int als_dataset_train_rows(const als_dataset *ds, const int **r) {
    if (!ds) return fail(ALS_E_INVALID, "null dataset");
    *r = ds->coo_row.data();
    return ALS_OK;
}
void als_dataset_free(als_dataset *ds) { delete ds; }

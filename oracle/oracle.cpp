/*
 * CPU oracle of AOCL-DA's distance path: a plain restatement of the reference algorithms,
 * used ONLY as the checker by tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py. Nothing in the product library links or calls it.
 *
 * Every function cites the reference file:line it restates. All arithmetic is double.
 * Pinning (see tests/test_oracle.py):
 *   - oracle_pairwise: FixedData_XY / FixedData_XX known answers of
 *     tests/unit_tests/metrics/metric_public.cpp:297-459, scipy.spatial.distance.cdist
 *     for every metric, and the invariants D(x,x)=0 and symmetry.
 *   - oracle_knn: the 6x3 example of nearest_neighbors_tests.hpp:97-144, the CSV fixtures
 *     tests/data/nn_data/*_5.csv (2092x5/2071 and 2125x20/2068, k=5), and
 *     sklearn.neighbors.NearestNeighbors(algorithm='brute').
 *   - oracle_kmeans: the supplied-centre cases of
 *     tests/unit_tests/kmeans/kmeans_test_data.hpp (3 clusters :250-308, large :444-528,
 *     empty clusters :141-193 incl. the split variant :789-805) and
 *     sklearn.cluster.KMeans(init=C, n_init=1, algorithm='lloyd') where no cluster
 *     empties; Elkan by the reference's Elkan cases (:600-603) and by equality with Lloyd's
 *     iterates where no two centres are equidistant; mixed-precision refinement by the
 *     reference's iteration counts (:944-969) and an IEEE rounding check of round_lp.
 *   - the random initialisations (k-means++ kmeans.cpp:1524-1664, AFK-MC2 :1667-1811,
 *     random :1404-1425, random partitions :1426-1457): the reference's inertia bounds of
 *     its random-initialisation cases (:604-716) and its fixed answers on all-zero data and
 *     the 1 x 1 matrix (:195-248,718-745); the k-means++ and AFK-MC2 draws are pinned
 *     against an independent restatement over libstdc++'s engine and distributions on
 *     integer data (tests/_stdrng.py), where the reference's GEMM-form weights are exact
 *     too. On general data the reference's weights carry BLAS rounding, so a draw that
 *     falls within an ulp of a boundary of the cumulative weights can pick another sample.
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <random>
#include <vector>

extern "C" {

/* Metric codes follow da_metric (aoclda_metrics.h:44-56). */
enum { M_EUCLID = 0, M_SQEUCLID = 1, M_MINKOWSKI = 2, M_MANHATTAN = 3, M_COSINE = 4,
       M_EUCLID_GEMM = 5, M_SQEUCLID_GEMM = 6, M_INNER = 7 };

static inline double at(int order, const double *A, int ld, int i, int j) {
    return order == 0 ? A[(size_t)i * ld + j] : A[i + (size_t)j * ld];
}

/*
 * da_pairwise_distances restated: metric switch pairwise_distances.cpp:36-75 (minkowski
 * p==2 -> euclidean, p==1 -> manhattan); direct forms of metrics_kernels.hpp:413-1070
 * ((x-y)^2, |x-y|, |x-y|^p, x*y); sqrt post-op euclidean_distance.cpp:74-83; cosine
 * cosine_distance.cpp:90-161 (norm <= eps -> 1, zero diagonal when Y is null); minkowski
 * pow(., 1/p); GEMM form euclidean_gemm_distance.cpp:59-230 (no clamp before sqrt, zero
 * diagonal and symmetric copy when Y is null); inner product = X Y^T.
 * eps is the machine epsilon of the caller's precision (cosine norm guard).
 */
int oracle_pairwise(int order, int m, int n, int k, const double *X, int ldx, const double *Y,
                    int ldy, double *D, int ldd, double p, int metric, double eps) {
    bool x_is_y = (Y == nullptr);
    if (x_is_y) {
        Y = X;
        n = m;
        ldy = ldx;
    }
    if (metric == M_MINKOWSKI && p == 2.0)
        metric = M_EUCLID;
    else if (metric == M_MINKOWSKI && p == 1.0)
        metric = M_MANHATTAN;
    std::vector<double> nx(m), ny(n);
    for (int i = 0; i < m; i++) {
        double s = 0;
        for (int l = 0; l < k; l++)
            s += at(order, X, ldx, i, l) * at(order, X, ldx, i, l);
        nx[i] = s;
    }
    for (int j = 0; j < n; j++) {
        double s = 0;
        for (int l = 0; l < k; l++)
            s += at(order, Y, ldy, j, l) * at(order, Y, ldy, j, l);
        ny[j] = s;
    }
    for (int i = 0; i < m; i++)
        for (int j = 0; j < n; j++) {
            double acc = 0;
            for (int l = 0; l < k; l++) {
                const double x = at(order, X, ldx, i, l), y = at(order, Y, ldy, j, l);
                switch (metric) {
                case M_EUCLID:
                case M_SQEUCLID:
                    acc += (x - y) * (x - y);
                    break;
                case M_MANHATTAN:
                    acc += std::fabs(x - y);
                    break;
                case M_MINKOWSKI:
                    acc += std::pow(std::fabs(x - y), p);
                    break;
                default:
                    acc += x * y;
                }
            }
            double v = acc;
            switch (metric) {
            case M_EUCLID:
                v = std::sqrt(acc);
                break;
            case M_MINKOWSKI:
                v = std::pow(acc, 1.0 / p);
                break;
            case M_COSINE: {
                double a = std::sqrt(nx[i]), b = std::sqrt(ny[j]);
                if (a <= eps)
                    a = 1.0;
                if (b <= eps)
                    b = 1.0;
                v = 1.0 - acc / (a * b);
                if (x_is_y && i == j)
                    v = 0.0;
                break;
            }
            case M_EUCLID_GEMM:
            case M_SQEUCLID_GEMM:
                v = nx[i] + ny[j] - 2.0 * acc;
                if (x_is_y && i == j)
                    v = 0.0;
                if (metric == M_EUCLID_GEMM)
                    v = std::sqrt(v);
                break;
            default:
                break;
            }
            if (order == 0)
                D[(size_t)i * ldd + j] = v;
            else
                D[i + (size_t)j * ldd] = v;
        }
    return 0;
}

/*
 * Brute-force kNN restated (nearest_neighbors.cpp:567-816 with set_params :295-352): the
 * euclidean metric is searched on squared distances and square-rooted at the end
 * (sorted_n_dist_n_ind, nearest_neighbors_utils.hpp:72-101); inner product keeps the k
 * largest. Ordering is by (key, index), which resolves the reference's schedule-dependent
 * tie order deterministically. Inputs are row-major; outputs row-major nq x k.
 */
int oracle_knn(int n, int d, const double *Xt, int nq, const double *Xq, int k, int metric,
               double p, int *ind, double *dist, double eps) {
    std::vector<double> keys(n);
    std::vector<int> perm(n);
    int km = metric;
    bool sq = false;
    if (metric == M_EUCLID || metric == M_EUCLID_GEMM || (metric == M_MINKOWSKI && p == 2.0)) {
        km = M_SQEUCLID;
        sq = true;
    } else if (metric == M_SQEUCLID_GEMM) {
        km = M_SQEUCLID;
    } else if (metric == M_MINKOWSKI && p == 1.0) {
        km = M_MANHATTAN;
    }
    for (int q = 0; q < nq; q++) {
        const double *x = Xq + (size_t)q * d;
        double nxq = 0;
        for (int l = 0; l < d; l++)
            nxq += x[l] * x[l];
        for (int j = 0; j < n; j++) {
            const double *y = Xt + (size_t)j * d;
            double acc = 0, ny = 0;
            for (int l = 0; l < d; l++) {
                switch (km) {
                case M_SQEUCLID:
                    acc += (x[l] - y[l]) * (x[l] - y[l]);
                    break;
                case M_MANHATTAN:
                    acc += std::fabs(x[l] - y[l]);
                    break;
                case M_MINKOWSKI:
                    acc += std::pow(std::fabs(x[l] - y[l]), p);
                    break;
                default:
                    acc += x[l] * y[l];
                    ny += y[l] * y[l];
                }
            }
            if (km == M_COSINE) {
                double a = std::sqrt(nxq), b = std::sqrt(ny);
                if (a <= eps)
                    a = 1.0;
                if (b <= eps)
                    b = 1.0;
                acc = 1.0 - acc / (a * b);
            } else if (km == M_INNER) {
                acc = -acc;
            }
            keys[j] = acc;
        }
        std::iota(perm.begin(), perm.end(), 0);
        std::partial_sort(perm.begin(), perm.begin() + k, perm.end(), [&](int a, int b) {
            return keys[a] < keys[b] || (keys[a] == keys[b] && a < b);
        });
        for (int j = 0; j < k; j++) {
            const int t = perm[j];
            double v = keys[t];
            if (sq)
                v = std::sqrt(v);
            else if (km == M_MINKOWSKI)
                v = std::pow(v, 1.0 / p);
            else if (km == M_INNER)
                v = -v;
            ind[(size_t)q * k + j] = t;
            if (dist)
                dist[(size_t)q * k + j] = v;
        }
    }
    return 0;
}

/* ---------------------------------------------------------------------------------------
 * k-means, restated from source/core/clustering/kmeans/kmeans.cpp, kmeans_lloyd.hpp and
 * kmeans_elkan.hpp. Data row-major n x d, centres row-major k x d, all arithmetic double.
 */
struct KMOut {
    int n_iter;
    int converged;
    double inertia;
    int empty_found;
};

static double sqd(const double *x, const double *c, int d) {
    double s = 0;
    for (int l = 0; l < d; l++)
        s += (x[l] - c[l]) * (x[l] - c[l]);
    return s;
}

/* Rounding to the lower precision of the mixed-precision phase (kmeans.cpp:1824-1876):
 * mode 1 rounds to float (the phase of double data), mode 2 to IEEE half through float
 * (the phase of float data), both to nearest even. */
static double round_lp(double v, int mode) {
    if (mode == 0)
        return v;
    const float f = (float)v;
    if (mode == 1 || !std::isfinite(f))
        return f;
    const double a = std::fabs((double)f);
    double q;
    if (a >= 65520.0)
        q = std::numeric_limits<double>::infinity();
    else if (a < 6.103515625e-05)              // below the smallest normal half: quantum 2^-24
        q = std::nearbyint(a * 16777216.0) / 16777216.0;
    else {
        int e;
        std::frexp(a, &e);                     // a in [2^(e-1), 2^e): 11 significant bits
        const double quantum = std::ldexp(1.0, e - 11);
        q = std::nearbyint(a / quantum) * quantum;
    }
    return std::copysign(q, (double)f);
}

double oracle_round_lp(double v, int mode) { return round_lp(v, mode); }

/* Spherical k-means (the "distance" option "cosine", kmeans.cpp:400-405): the inverse data
 * norms safe_inv(|x|) when the data are normalised (kmeans.cpp:681-703), else none. A null
 * Sph pointer below means the Euclidean k-means. */
struct Sph {
    int normalize;
    std::vector<double> inv;
};

static double dotp(const double *a, const double *b, int d) {
    double s = 0;
    for (int l = 0; l < d; l++)
        s += a[l] * b[l];
    return s;
}

/* normalize_rows_inplace, da_utils.cpp:807-855: a zero row is left as it is. */
static void normalize_rows(int k, int d, double *C) {
    for (int c = 0; c < k; c++) {
        const double nrm = std::sqrt(dotp(C + (size_t)c * d, C + (size_t)c * d, d));
        const double inv = nrm == 0 ? 1.0 : 1.0 / nrm;
        for (int l = 0; l < d; l++)
            C[(size_t)c * d + l] *= inv;
    }
}

/* The distance of sample i to a unit centre c: 1 - x.c/|x| (or 1 - x.c for data taken as
 * normalised), kmeans.cpp:1171-1199; squared Euclidean without g. */
static double centre_dist(int d, const double *X, const Sph *g, int i, const double *c) {
    if (!g)
        return sqd(X + (size_t)i * d, c, d);
    double cs = dotp(X + (size_t)i * d, c, d);
    if (g->normalize)
        cs *= g->inv[i];
    return 1.0 - cs;
}

/* compute_distances_to_point, kmeans.cpp:1488-1522: the distance of sample i to sample j,
 * for spherical k-means max(0, 1 - cos). */
static double sample_dist(int d, const double *X, const Sph *g, int i, int j) {
    if (!g)
        return sqd(X + (size_t)i * d, X + (size_t)j * d, d);
    double cs = dotp(X + (size_t)i * d, X + (size_t)j * d, d);
    if (g->normalize)
        cs *= g->inv[i] * g->inv[j];
    return std::max(0.0, 1.0 - cs);
}

/* lloyd_iteration (kmeans_lloyd.hpp:105-271): labels by argmin with ties to the lowest
 * centre (lloyd_iteration_kernel_scalar, kmeans_kernels.cpp:209-260); with update, the sums
 * per label and scale_current_cluster_centres :363-405 (empty -> count 1 -> origin).
 * Returns the number of changed labels. */
static int lloyd_iter(int n, int d, int k, const double *X, const double *Cprev, double *Cnew,
                      int *lab, const int *lab_prev, bool update, const Sph *g = nullptr) {
    int changed = 0;
    for (int i = 0; i < n; i++) {
        double best = std::numeric_limits<double>::infinity();
        int bl = 0;
        for (int c = 0; c < k; c++) {
            // spherical: the GEMM form -x.c of kmeans_lloyd.hpp:144-146 (centre norms zero)
            const double v = g ? -dotp(X + (size_t)i * d, Cprev + (size_t)c * d, d)
                               : sqd(X + (size_t)i * d, Cprev + (size_t)c * d, d);
            if (v < best) {
                best = v;
                bl = c;
            }
        }
        if (lab_prev && lab_prev[i] != bl)
            changed++;
        lab[i] = bl;
    }
    if (update) {
        std::vector<int> cnt(k, 0);
        std::fill(Cnew, Cnew + (size_t)k * d, 0.0);
        for (int i = 0; i < n; i++) {
            cnt[lab[i]]++;
            // normalised spherical data: sums of x/|x| (kmeans_lloyd.hpp:279-300)
            const double w = (g && g->normalize) ? g->inv[i] : 1.0;
            for (int l = 0; l < d; l++)
                Cnew[(size_t)lab[i] * d + l] += X[(size_t)i * d + l] * w;
        }
        if (g) {
            normalize_rows(k, d, Cnew);
        } else {
            for (int c = 0; c < k; c++)
                for (int l = 0; l < d; l++)
                    Cnew[(size_t)c * d + l] /= (cnt[c] == 0 ? 1 : cnt[c]);
        }
    }
    return changed;
}

/* Elkan's bounds (kmeans_elkan.hpp:143-566, kmeans_kernels.cpp:116-135). u: upper bound of
 * each sample's distance to its centre; l: n x k lower bounds; cc: half centre-to-centre
 * distances; nxt: each centre's half distance to its nearest other centre. */
struct Elkan {
    std::vector<double> u, l, cc, nxt;
};

/* compute_centre_half_distances, kmeans_elkan.hpp:515-566 (distances, not squares). */
static void elkan_half_distances(int k, int d, const double *C, Elkan &e) {
    e.cc.assign((size_t)k * k, 0.0);
    e.nxt.assign(k, std::numeric_limits<double>::infinity());
    for (int j = 0; j < k; j++)
        for (int i = 0; i < j; i++) {
            const double t = 0.5 * std::sqrt(sqd(C + (size_t)i * d, C + (size_t)j * d, d));
            e.cc[(size_t)i * k + j] = e.cc[(size_t)j * k + i] = t;
            e.nxt[i] = std::min(e.nxt[i], t);
            e.nxt[j] = std::min(e.nxt[j], t);
        }
}

/* init_elkan, kmeans_elkan.hpp:143-233: centre 0 first, a centre j is measured only when the
 * best distance so far exceeds half the distance between the best centre and j; strict <. */
static void elkan_init(int n, int d, int k, const double *X, const double *C, int *lab,
                       Elkan &e) {
    elkan_half_distances(k, d, C, e);
    e.u.assign(n, 0.0);
    e.l.assign((size_t)n * k, 0.0);
    for (int i = 0; i < n; i++) {
        const double *x = X + (size_t)i * d;
        int label = 0;
        double sd = std::sqrt(sqd(x, C, d));
        e.l[(size_t)i * k] = sd;
        for (int j = 1; j < k; j++) {
            e.l[(size_t)i * k + j] = 0.0;
            if (sd > e.cc[(size_t)label * k + j]) {
                const double dist = std::sqrt(sqd(x, C + (size_t)j * d, d));
                e.l[(size_t)i * k + j] = dist;
                if (dist < sd) {
                    label = j;
                    sd = dist;
                }
            }
        }
        lab[i] = label;
        e.u[i] = sd;
    }
}

/* elkan_iteration_assign_block, kmeans_elkan.hpp:427-513: samples whose upper bound exceeds
 * the half distance to the next centre test each other centre j against its lower bound and
 * the half distance; the first such test tightens the upper bound; a measured distance
 * replaces the label only when strictly smaller. With update, the sums per label as in
 * Lloyd's. */
static int elkan_assign(int n, int d, int k, const double *X, const double *Cprev,
                        double *Cnew, int *lab, const int *lab_prev, bool update, Elkan &e) {
    int changed = 0;
    for (int i = 0; i < n; i++) {
        const double *x = X + (size_t)i * d;
        double *lb = e.l.data() + (size_t)i * k;
        int label = lab_prev[i];
        double u = e.u[i];
        bool tight = false;
        if (u > e.nxt[label]) {
            for (int j = 0; j < k; j++) {
                const double l_j = lb[j];
                const double h = e.cc[(size_t)label * k + j];
                if (j != label && u > l_j && u > h) {
                    if (!tight) {
                        u = std::sqrt(sqd(x, Cprev + (size_t)label * d, d));
                        lb[label] = u;
                        tight = true;
                    }
                    if (u > l_j || u > h) {
                        const double dist = std::sqrt(sqd(x, Cprev + (size_t)j * d, d));
                        lb[j] = dist;
                        if (dist < u) {
                            u = dist;
                            label = j;
                        }
                    }
                }
            }
        }
        e.u[i] = u;
        if (label != lab_prev[i])
            changed++;
        lab[i] = label;
    }
    if (update) {
        std::vector<int> cnt(k, 0);
        std::fill(Cnew, Cnew + (size_t)k * d, 0.0);
        for (int i = 0; i < n; i++) {
            cnt[lab[i]]++;
            for (int l = 0; l < d; l++)
                Cnew[(size_t)lab[i] * d + l] += X[(size_t)i * d + l];
        }
        for (int c = 0; c < k; c++)
            for (int l = 0; l < d; l++)
                Cnew[(size_t)c * d + l] /= (cnt[c] == 0 ? 1 : cnt[c]);
    }
    return changed;
}

/* The bound update after the centres moved (elkan_iteration, kmeans_elkan.hpp:372-418, and
 * elkan_iteration_kernel_scalar, kmeans_kernels.cpp:116-131): u += shift of the sample's
 * centre, l_j = max(l_j - shift_j, 0), then the half distances of the new centres. */
static void elkan_update(int n, int d, int k, const double *Cprev, const double *Cnew,
                         const int *lab, Elkan &e) {
    std::vector<double> shift(k);
    for (int c = 0; c < k; c++)
        shift[c] = std::sqrt(sqd(Cprev + (size_t)c * d, Cnew + (size_t)c * d, d));
    for (int i = 0; i < n; i++) {
        e.u[i] += shift[lab[i]];
        for (int j = 0; j < k; j++) {
            double &v = e.l[(size_t)i * k + j];
            v -= shift[j];
            if (v < 0)
                v = 0.0;
        }
    }
    elkan_half_distances(k, d, Cnew, e);
}

/* handle_empty_clusters, kmeans.cpp:1127-1330 (split mode in double arithmetic). Returns
 * 1 if a split was made, 2 if empty clusters exist in error mode, 0 otherwise. */
static int empty_clusters(int n, int d, int k, const double *X, double *C, int *lab, int mode,
                          const Sph *g = nullptr) {
    std::vector<int> cnt(k, 0);
    for (int i = 0; i < n; i++)
        cnt[lab[i]]++;
    int n_empty = 0;
    for (int c = 0; c < k; c++)
        n_empty += cnt[c] == 0;
    if (n_empty == 0)
        return 0;
    if (mode == 1)
        return 2;
    std::vector<double> dist(n);
    for (int i = 0; i < n; i++)
        dist[i] = centre_dist(d, X, g, i, C + (size_t)lab[i] * d);
    for (int c = 0; c < k; c++) {
        if (cnt[c] != 0)
            continue;
        int best = -1;
        double bd = -1;
        for (int i = 0; i < n; i++)
            if (cnt[lab[i]] > 1 && std::isfinite(dist[i]) && dist[i] > bd) {
                bd = dist[i];
                best = i;
            }
        if (best < 0)
            break;
        const int old = lab[best], n_old = cnt[old];
        const double wb = (g && g->normalize) ? g->inv[best] : 1.0;
        for (int l = 0; l < d; l++) {
            double &cv = C[(size_t)old * d + l];
            cv = (cv * n_old - X[(size_t)best * d + l] * wb) / (n_old - 1);
        }
        for (int l = 0; l < d; l++)
            C[(size_t)c * d + l] = X[(size_t)best * d + l] * wb;
        if (g) {
            normalize_rows(1, d, C + (size_t)old * d);
            normalize_rows(1, d, C + (size_t)c * d);
        }
        cnt[old]--;
        cnt[c] = 1;
        lab[best] = c;
        dist[best] = 0;
        if (n_empty > 1)
            for (int i = 0; i < n; i++)
                if (lab[i] == old)
                    dist[i] = centre_dist(d, X, g, i, C + (size_t)old * d);
    }
    return 1;
}

/* perform_kmeans, kmeans.cpp:996-1058, convergence_test :1358-1397 and
 * compute_current_inertia :1060-1124, for Lloyd (algorithm 0) or Elkan (1). C holds the
 * initial centres on entry and the final centres on exit. lp_mode != 0 runs the phase of
 * mixed-precision refinement in the lower precision (kmeans.cpp:1824-1876): X is already
 * rounded, and the centres are rounded after every update, as kmeans<float> or
 * kmeans<_Float16> stores them. */
static KMOut perform(int n, int d, int k, const double *X, double *C, int *lab, int max_iter,
                     double tol, int empty_mode, int algorithm, int lp_mode,
                     const Sph *g = nullptr) {
    std::vector<double> Cp((size_t)k * d);
    std::vector<int> lp(n, 0);
    std::vector<double> Cc(C, C + (size_t)k * d);
    std::vector<int> lc(n, 0);
    KMOut o{0, 0, 0.0, 0};
    const bool elkan = algorithm == 1;
    Elkan e;
    if (elkan)
        elkan_init(n, d, k, X, Cc.data(), lc.data(), e);
    int conv = 0, it;
    for (it = 0; it < max_iter; it++) {
        std::swap(Cp, Cc);
        std::swap(lp, lc);
        const int changed =
            elkan ? elkan_assign(n, d, k, X, Cp.data(), Cc.data(), lc.data(), lp.data(), true, e)
                  : lloyd_iter(n, d, k, X, Cp.data(), Cc.data(), lc.data(), lp.data(), true, g);
        for (auto &v : Cc)
            v = round_lp(v, lp_mode);
        // compute_centre_shift, kmeans.cpp:1334-1355 (Frobenius norms).
        double normc2 = 0, shift2 = 0;
        for (size_t t = 0; t < (size_t)k * d; t++) {
            normc2 += Cp[t] * Cp[t];
            shift2 += (Cp[t] - Cc[t]) * (Cp[t] - Cc[t]);
        }
        if (elkan)
            elkan_update(n, d, k, Cp.data(), Cc.data(), lc.data(), e);
        if (empty_mode != 0) {
            const int r = empty_clusters(n, d, k, X, Cc.data(), lc.data(), empty_mode, g);
            if (r == 2) {
                o.empty_found = 1;
                return o;
            }
            if (r == 1) {
                if (elkan) {
                    // Bounds invalidated after a split (kmeans.cpp:1312-1317).
                    std::fill(e.u.begin(), e.u.end(), std::numeric_limits<double>::max());
                    std::fill(e.l.begin(), e.l.end(), 0.0);
                }
                continue;
            }
        }
        conv = 0;
        if (it > 1 && changed == 0)
            conv = 2;
        else if (std::sqrt(shift2) < tol * std::sqrt(normc2))
            conv = 1;
        if (conv)
            break;
    }
    o.n_iter = it;
    o.converged = conv;
    if (conv == 1 || it == max_iter) {
        std::swap(lp, lc);
        std::swap(Cp, Cc);
        if (elkan)
            elkan_assign(n, d, k, X, Cp.data(), nullptr, lc.data(), lp.data(), false, e);
        else
            lloyd_iter(n, d, k, X, Cp.data(), nullptr, lc.data(), nullptr, false, g);
        std::swap(Cp, Cc);
    }
    double inertia = 0;
    for (int i = 0; i < n; i++)
        inertia += centre_dist(d, X, g, i, Cc.data() + (size_t)lc[i] * d);
    o.inertia = inertia;
    std::copy(Cc.begin(), Cc.end(), C);
    std::copy(lc.begin(), lc.end(), lab);
    return o;
}

/* kmeans_plusplus, kmeans.cpp:1524-1664, with exact squared distances (spherical: 1 - cos,
 * clamped at zero) as weights. */
static void kmeanspp(int n, int d, int k, const double *X, double *C, std::mt19937_64 &mt,
                     const Sph *g = nullptr) {
    const int n_trials = 2 + (int)std::log((double)k);
    std::uniform_int_distribution<> dis_int(0, n - 1);
    const int first = dis_int(mt);
    std::vector<double> w(n), w2(n), w5(n), w4(n);
    for (int i = 0; i < n; i++) {
        double v = sample_dist(d, X, g, i, first);
        w[i] = (!std::isfinite(v) || v < 0) ? 0.0 : v;
    }
    w[first] = 0;
    std::vector<int> chosen(k, 0);
    chosen[0] = first;
    bool coincident = true;
    for (int i = 0; i < n; i++)
        if (w[i] > 0) {
            coincident = false;
            break;
        }
    if (coincident) {
        for (int c = 0; c < k; c++)
            chosen[c] = c;
    } else {
        std::vector<int> cand(n_trials);
        for (int c = 1; c < k; c++) {
            int best = 0;
            double best_cost = std::numeric_limits<double>::infinity();
            std::discrete_distribution<> dd(w.begin(), w.end());
            for (int t = 0; t < n_trials; t++)
                cand[t] = dd(mt);
            for (int t = 0; t < n_trials; t++) {
                bool rep = false;
                for (int j = 0; j < t; j++)
                    rep |= cand[j] == cand[t];
                if (rep)
                    continue;
                double cost = 0;
                for (int i = 0; i < n; i++) {
                    w4[i] = sample_dist(d, X, g, i, cand[t]);
                    w5[i] = std::max(0.0, std::min(w[i], w4[i]));
                    cost += w5[i];
                }
                if (std::isfinite(cost) && cost < best_cost) {
                    best_cost = cost;
                    best = cand[t];
                    std::swap(w2, w5);
                }
            }
            chosen[c] = best;
            std::swap(w, w2);
            w[best] = 0;
        }
    }
    for (int c = 0; c < k; c++)
        for (int l = 0; l < d; l++)
            C[(size_t)c * d + l] = X[(size_t)chosen[c] * d + l];
}

/* afk_mcmc_init, kmeans.cpp:1667-1811 (Bachem et al. 2016): the first centre uniform; the
 * proposal q(x) = d(x,c1)^2 / (2 sum) + 1/(2n) (uniform when every distance is zero); for
 * each further centre a Markov chain of m steps over draws from q, accepting y with
 * probability min(1, d(y)q(x) / (d(x)q(y))), where d is the squared distance to the
 * nearest centre chosen so far; a zero d(x)q(y) accepts without a uniform draw. Spherical
 * k-means: the distances max(0, 1 - cos) (the proposal between samples, the chain against the
 * centres stored as x/|x|, or as x for data taken as unit rows). */
static void afkmc2(int n, int d, int k, const double *X, double *C, int m, std::mt19937_64 &mt,
                   const Sph *g = nullptr) {
    std::uniform_int_distribution<> dis_int(0, n - 1);
    const int c1 = dis_int(mt);
    auto place = [&](int c, int idx) {
        const double wv = (g && g->normalize) ? g->inv[idx] : 1.0;
        for (int l = 0; l < d; l++)
            C[(size_t)c * d + l] = X[(size_t)idx * d + l] * wv;
    };
    place(0, c1);
    std::vector<double> w(n), q(n);
    double sum = 0;
    for (int i = 0; i < n; i++) {
        w[i] = std::max(0.0, sample_dist(d, X, g, i, c1));
        sum += w[i];
    }
    const double inv_2n = 1.0 / (2.0 * n);
    for (int i = 0; i < n; i++)
        q[i] = (std::isfinite(sum) && sum > 0) ? 0.5 * w[i] / sum + inv_2n : 1.0 / n;
    auto nearest = [&](int idx, int nc) {
        double best = std::numeric_limits<double>::infinity();
        for (int c = 0; c < nc; c++)
            best = std::min(best, g ? std::max(0.0, centre_dist(d, X, g, idx, C + (size_t)c * d))
                                    : sqd(X + (size_t)idx * d, C + (size_t)c * d, d));
        return best;
    };
    std::discrete_distribution<> qd(q.begin(), q.end());
    std::uniform_real_distribution<double> unif(0.0, 1.0);
    for (int i = 1; i < k; i++) {
        int x = qd(mt);
        double dx = nearest(x, i);
        for (int s = 0; s < m; s++) {
            const int y = qd(mt);
            const double dy = nearest(y, i);
            const double dxqy = dx * q[y];
            if (dxqy == 0.0 || (dy * q[x]) / dxqy > unif(mt)) {
                x = y;
                dx = dy;
            }
        }
        place(i, x);
    }
}

/* The initial centres of the first run (k-means++ init 1, AFK-MC2 init 4) for the given
 * seed, as oracle_kmeans draws them; exposed so that tests can check the draws. */
int oracle_kmeans_init(int n, int d, const double *X, int k, int init, long long seed,
                       int afk_m, double *C_out) {
    std::mt19937_64 mt((unsigned long long)seed);
    if (init == 1)
        kmeanspp(n, d, k, X, C_out, mt);
    else if (init == 4)
        afkmc2(n, d, k, X, C_out, afk_m, mt);
    else
        return -1;
    return 0;
}

/*
 * kmeans<T>::compute restated (kmeans.cpp:351-783) for Lloyd (algorithm 0) or Elkan (1)
 * with initialisation 0 random (da_std::sample, :1404-1425), 1 k-means++, 2 supplied (C_init),
 * 3 random partitions (:1426-1457), 4 AFK-MC2 with afk_m chain steps. empty_mode: 0
 * ignore, 1 error, 2 split. lp_mode: 0 none, 1 mixed precision of double data (float
 * phase), 2 of float data (half phase), with lp_tol and lp_max_iter. Returns 0 on success,
 * 900 when every run hit an empty cluster (error mode), 21 when the best run hit max_iter,
 * 41 when lp_tol <= tol (incompatible options). out_info = {n_iter, inertia, lp_n_iter}.
 * Centres row-major k x d.
 */
static int kmeans_impl(int n, int d, const double *X, int k, int init, const double *C_init,
                       int n_init, int max_iter, double tol, long long seed, int empty_mode,
                       int algorithm, int afk_m, int lp_mode, double lp_tol, int lp_max_iter,
                       double *C_out, int *labels_out, double *out_info, const Sph *g) {
    if (n_init > 1 && init == 2)
        n_init = 1;
    if (lp_mode != 0 && lp_tol <= tol)
        return 41;
    std::mt19937_64 mt;
    if (init != 2)
        mt.seed((unsigned long long)seed);
    std::vector<double> C((size_t)k * d);
    std::vector<int> lab(n);
    std::vector<double> Xlp;
    if (lp_mode != 0) {
        Xlp.resize((size_t)n * d);
        for (size_t t = 0; t < Xlp.size(); t++)
            Xlp[t] = round_lp(X[t], lp_mode);
    }
    double best = std::numeric_limits<double>::infinity();
    bool valid = false, warn = false;
    for (int run = 0; run < n_init; run++) {
        if (init == 2) {
            std::copy(C_init, C_init + (size_t)k * d, C.begin());
        } else if (init == 0) {
            // da_std::sample, da_std.hpp:102-118 (selection sampling, one draw per sample).
            std::vector<int> pick;
            long long need = k, left = n;
            for (int i = 0; need > 0; i++, left--) {
                std::uniform_int_distribution<long long> u(0, left - 1);
                if (u(mt) < need) {
                    pick.push_back(i);
                    need--;
                }
            }
            for (int c = 0; c < k; c++)
                for (int l = 0; l < d; l++)
                    C[(size_t)c * d + l] = X[(size_t)pick[c] * d + l];
        } else if (init == 3) {
            std::uniform_int_distribution<> dis(0, k - 1);
            std::vector<int> cnt(k, 0);
            std::fill(C.begin(), C.end(), 0.0);
            for (int i = 0; i < n; i++) {
                const int c = dis(mt);
                cnt[c]++;
                for (int l = 0; l < d; l++)
                    C[(size_t)c * d + l] += X[(size_t)i * d + l];
            }
            for (int c = 0; c < k; c++)
                for (int l = 0; l < d; l++)
                    C[(size_t)c * d + l] /= cnt[c];
        } else if (init == 4) {
            afkmc2(n, d, k, X, C.data(), afk_m, mt, g);
        } else {
            kmeanspp(n, d, k, X, C.data(), mt, g);
        }
        // spherical: the initial centres normalised (kmeans.cpp:1478-1486)
        if (g)
            normalize_rows(k, d, C.data());
        int lp_iter = 0;
        if (lp_mode != 0) {
            for (auto &v : C)
                v = round_lp(v, lp_mode);
            std::vector<int> lab_lp(n);
            KMOut lo = perform(n, d, k, Xlp.data(), C.data(), lab_lp.data(), lp_max_iter, lp_tol,
                               empty_mode, algorithm, lp_mode);
            if (lo.empty_found)
                return 900;
            lp_iter = lo.n_iter;
            for (auto &v : C)
                v = round_lp(v, lp_mode);
        }
        KMOut o = perform(n, d, k, X, C.data(), lab.data(), max_iter, tol, empty_mode, algorithm,
                          0, g);
        if (o.empty_found)
            continue;
        valid = true;
        if (o.inertia < best || run == 0 || !std::isfinite(best)) {
            best = o.inertia;
            warn = o.converged == 0;
            std::copy(C.begin(), C.end(), C_out);
            std::copy(lab.begin(), lab.end(), labels_out);
            out_info[0] = o.n_iter;
            out_info[1] = o.inertia;
            out_info[2] = lp_iter;
        }
    }
    if (!valid)
        return 900;
    return warn ? 21 : 0;
}

int oracle_kmeans(int n, int d, const double *X, int k, int init, const double *C_init,
                  int n_init, int max_iter, double tol, long long seed, int empty_mode,
                  int algorithm, int afk_m, int lp_mode, double lp_tol, int lp_max_iter,
                  double *C_out, int *labels_out, double *out_info) {
    return kmeans_impl(n, d, X, k, init, C_init, n_init, max_iter, tol, seed, empty_mode,
                       algorithm, afk_m, lp_mode, lp_tol, lp_max_iter, C_out, labels_out,
                       out_info, nullptr);
}

/* Spherical k-means (the "distance" option "cosine"), Lloyd, full precision, initialisation
 * 0 random, 1 k-means++, 2 supplied, 3 random partitions or 4 AFK-MC2 with afk_m chain steps
 * (kmeans.cpp:1667-1811), empty_mode 0 ignore, 1 error or
 * 2 split (kmeans.cpp:1160-1320, the relocated sample and the donor centre normalised), with
 * the data normalised (normalize 1, kmeans.cpp:681-703) or taken as unit rows (0): labels by
 * the largest x.c (ties to the lowest centre), centres the normalised sums of the members (of
 * x/|x| when normalised), inertia the sum of 1 - x.c/|x| (1 - x.c). Returns as
 * oracle_kmeans; out_info = {n_iter, inertia, 0}. */
int oracle_kmeans_sph(int n, int d, const double *X, int k, int init, const double *C_init,
                      int n_init, int max_iter, double tol, long long seed, int empty_mode,
                      int afk_m, int normalize, double *C_out, int *labels_out,
                      double *out_info) {
    Sph g{normalize, {}};
    if (normalize) {
        g.inv.resize(n);
        for (int i = 0; i < n; i++) {
            const double nrm = std::sqrt(dotp(X + (size_t)i * d, X + (size_t)i * d, d));
            g.inv[i] = nrm != 0 ? 1.0 / nrm : 0.0;
        }
    }
    return kmeans_impl(n, d, X, k, init, C_init, n_init, max_iter, tol, seed, empty_mode, 0,
                       afk_m, 0, 0.0, 0, C_out, labels_out, out_info, &g);
}

/*
 * Brute-force radius neighbours restated (nearest_neighbors.cpp:1918-2080 and the brute-force
 * kernel :2080-2215, set_params :337-347, extraction :2310-2330): the euclidean forms
 * (euclidean, euclidean_gemm, minkowski p = 2) compare squared euclidean distances with
 * radius^2 and return square roots; sqeuclidean(_gemm), manhattan, cosine and minkowski
 * (p != 2, with its root) compare the distance itself with radius. Inclusive (<=). With
 * single != 0 the distance and the threshold are rounded to float before the comparison,
 * as the reference's single-precision path computes them in float. Neighbours of a query
 * are listed in ascending index (the single-threaded insertion order), or, with sort != 0,
 * stably sorted by distance (sorted_n_dist_n_ind). counts[q] receives the number of
 * neighbours of query q; ind/dist (capacity cap) the neighbours of all queries in query
 * order. Returns the total number of neighbours.
 */
long long oracle_radius(int n, int d, const double *Xt, int nq, const double *Xq, int metric,
                        double p, double r, int single, int sort, int *counts, int *ind,
                        double *dist, long long cap, double eps) {
    const bool sq = metric == M_EUCLID || metric == M_EUCLID_GEMM ||
                    (metric == M_MINKOWSKI && p == 2.0);
    int km = metric;
    if (sq)
        km = metric == M_EUCLID_GEMM ? M_SQEUCLID_GEMM : M_SQEUCLID;
    double thr = sq ? r * r : r;
    if (single) {
        const float rf = (float)r;
        thr = sq ? (double)(rf * rf) : (double)rf;
    }
    std::vector<double> D(n);
    std::vector<int> hit;
    long long total = 0;
    for (int q = 0; q < nq; q++) {
        oracle_pairwise(0, 1, n, d, Xq + (size_t)q * d, d, Xt, d, D.data(), n, p, km, eps);
        hit.clear();
        for (int j = 0; j < n; j++) {
            double v = D[j];
            if (single)
                v = (double)(float)v;
            D[j] = v;
            if (v <= thr)
                hit.push_back(j);
        }
        if (sort)
            std::stable_sort(hit.begin(), hit.end(), [&](int a, int b) { return D[a] < D[b]; });
        counts[q] = (int)hit.size();
        for (int j : hit) {
            if (total < cap) {
                ind[total] = j;
                double v = D[j];
                if (sq)
                    v = single ? (double)std::sqrt((float)v) : std::sqrt(v);
                dist[total] = v;
            }
            total++;
        }
    }
    return total;
}

/*
 * DBSCAN restated (dbscan.cpp:214-291 compute, :395-445 the serial cluster formation,
 * :449-470 core points; neighbour lists of radius_neighbors.cpp:47-200). Neighbours of i are
 * the j != i whose key is <= the threshold: metric 5 ("euclidean") compares squared
 * distances with eps^2; 0 ("l2") the distance with eps; 6 ("sqeuclidean") the squared
 * distance with eps; minkowski with p = 2 the squared distance with eps itself, p = 1 the
 * manhattan distance, other p the minkowski distance; manhattan and cosine the distance.
 * Keys are exact (direct differences in double); with single != 0 the key and the
 * threshold are rounded to float first, as a single-precision build computes them in float.
 * A point with at least min_samples - 1 neighbours is a core point. Clusters are grown
 * breadth first from each unvisited core point in index order; a non-core point reached
 * from a cluster joins it, otherwise it is noise (-1). Returns the number of clusters;
 * core[i] receives 1 for core points. Inputs row-major n x d.
 */
int oracle_dbscan(int n, int d, const double *X, int metric, double p, double eps,
                  int min_samples, int single, int *labels, int *core) {
    int km = metric;
    double thr = eps;
    if (metric == M_EUCLID_GEMM) {
        km = M_SQEUCLID;
        thr = single ? (double)((float)eps * (float)eps) : eps * eps;
    } else if (metric == M_SQEUCLID_GEMM) {
        km = M_SQEUCLID;
    } else if (metric == M_MINKOWSKI && p == 2.0) {
        km = M_SQEUCLID;
    } else if (metric == M_MINKOWSKI && p == 1.0) {
        km = M_MANHATTAN;
    }
    if (single)
        thr = (double)(float)thr;
    std::vector<std::vector<int>> nb(n);
    std::vector<double> D(n);
    for (int i = 0; i < n; i++) {
        oracle_pairwise(0, 1, n, d, X + (size_t)i * d, d, X, d, D.data(), n, p, km,
                        single ? 1.1920928955078125e-07 : 2.220446049250313e-16);
        for (int j = 0; j < n; j++) {
            double v = D[j];
            if (single)
                v = (double)(float)v;
            if (j != i && v <= thr)
                nb[i].push_back(j);
        }
    }
    for (int i = 0; i < n; i++)
        core[i] = (int)nb[i].size() >= min_samples - 1 ? 1 : 0;
    const int UNVISITED = -2, NOISE = -1;
    for (int i = 0; i < n; i++)
        labels[i] = UNVISITED;
    int nc = 0;
    for (int i = 0; i < n; i++) {
        if (labels[i] != UNVISITED)
            continue;
        if (!core[i]) {
            labels[i] = NOISE;
            continue;
        }
        labels[i] = nc;
        std::vector<int> search(nb[i]);
        for (size_t t = 0; t < search.size(); t++) {
            const int j = search[t];
            if (labels[j] == NOISE) {
                labels[j] = nc;
                continue;
            }
            if (labels[j] != UNVISITED)
                continue;
            labels[j] = nc;
            if (core[j])
                search.insert(search.end(), nb[j].begin(), nb[j].end());
        }
        nc++;
    }
    return nc;
}

} // extern "C"

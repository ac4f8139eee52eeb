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
 *     empties. The k-means++ centre sequence (kmeans.cpp:1524-1664) is restated with the
 *     same libstdc++ engine and distributions; for it only the distribution of draws is
 *     pinned (inertia bound of kmeans_test_data.hpp:586-700 style tests), the exact
 *     sequence of the reference is "parity unpinned" because its weights carry BLAS
 *     rounding.
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
 * k-means (Lloyd; Elkan gives the same iterates), restated from
 * source/core/clustering/kmeans/kmeans.cpp and kmeans_lloyd.hpp. Data row-major n x d.
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

/* lloyd_iteration (kmeans_lloyd.hpp:105-271): labels by argmin with ties to the lowest
 * centre (lloyd_iteration_kernel_scalar, kmeans_kernels.cpp:209-260), sums per label,
 * scale_current_cluster_centres :363-405 (empty -> count 1 -> origin), shift
 * compute_centre_shift kmeans.cpp:1334-1355. Returns the number of changed labels. */
static int lloyd_iter(int n, int d, int k, const double *X, const double *Cprev, double *Cnew,
                      int *lab, const int *lab_prev, bool update, double *normc2,
                      double *shift2) {
    int changed = 0;
    for (int i = 0; i < n; i++) {
        double best = std::numeric_limits<double>::infinity();
        int bl = 0;
        for (int c = 0; c < k; c++) {
            const double v = sqd(X + (size_t)i * d, Cprev + (size_t)c * d, d);
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
            for (int l = 0; l < d; l++)
                Cnew[(size_t)lab[i] * d + l] += X[(size_t)i * d + l];
        }
        for (int c = 0; c < k; c++)
            for (int l = 0; l < d; l++)
                Cnew[(size_t)c * d + l] /= (cnt[c] == 0 ? 1 : cnt[c]);
        double a = 0, b = 0;
        for (size_t t = 0; t < (size_t)k * d; t++) {
            a += Cprev[t] * Cprev[t];
            b += (Cprev[t] - Cnew[t]) * (Cprev[t] - Cnew[t]);
        }
        *normc2 = a;
        *shift2 = b;
    }
    return changed;
}

/* handle_empty_clusters, kmeans.cpp:1127-1330 (split mode in double arithmetic). Returns
 * 1 if a split was made, 2 if empty clusters exist in error mode, 0 otherwise. */
static int empty_clusters(int n, int d, int k, const double *X, double *C, int *lab, int mode) {
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
        dist[i] = sqd(X + (size_t)i * d, C + (size_t)lab[i] * d, d);
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
        for (int l = 0; l < d; l++) {
            double &cv = C[(size_t)old * d + l];
            cv = (cv * n_old - X[(size_t)best * d + l]) / (n_old - 1);
        }
        for (int l = 0; l < d; l++)
            C[(size_t)c * d + l] = X[(size_t)best * d + l];
        cnt[old]--;
        cnt[c] = 1;
        lab[best] = c;
        dist[best] = 0;
        if (n_empty > 1)
            for (int i = 0; i < n; i++)
                if (lab[i] == old)
                    dist[i] = sqd(X + (size_t)i * d, C + (size_t)old * d, d);
    }
    return 1;
}

/* perform_kmeans, kmeans.cpp:996-1058, and compute_current_inertia :1060-1124. C holds the
 * initial centres on entry and the final centres on exit. */
static KMOut perform(int n, int d, int k, const double *X, double *C, int *lab, int max_iter,
                     double tol, int empty_mode) {
    std::vector<double> Cp((size_t)k * d);
    std::vector<int> lp(n, 0);
    std::vector<double> Cc(C, C + (size_t)k * d);
    std::vector<int> lc(n, 0);
    KMOut o{0, 0, 0.0, 0};
    int conv = 0, it;
    double normc2 = 0, shift2 = 0;
    for (it = 0; it < max_iter; it++) {
        std::swap(Cp, Cc);
        std::swap(lp, lc);
        const int changed = lloyd_iter(n, d, k, X, Cp.data(), Cc.data(), lc.data(), lp.data(), true,
                                       &normc2, &shift2);
        if (empty_mode != 0) {
            const int r = empty_clusters(n, d, k, X, Cc.data(), lc.data(), empty_mode);
            if (r == 2) {
                o.empty_found = 1;
                return o;
            }
            if (r == 1)
                continue;
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
        lloyd_iter(n, d, k, X, Cp.data(), nullptr, lc.data(), nullptr, false, nullptr, nullptr);
        std::swap(Cp, Cc);
    }
    double inertia = 0;
    for (int i = 0; i < n; i++)
        inertia += sqd(X + (size_t)i * d, Cc.data() + (size_t)lc[i] * d, d);
    o.inertia = inertia;
    std::copy(Cc.begin(), Cc.end(), C);
    std::copy(lc.begin(), lc.end(), lab);
    return o;
}

/* kmeans_plusplus, kmeans.cpp:1524-1664, with exact squared distances as weights. */
static void kmeanspp(int n, int d, int k, const double *X, double *C, std::mt19937_64 &mt) {
    const int n_trials = 2 + (int)std::log((double)k);
    std::uniform_int_distribution<> dis_int(0, n - 1);
    const int first = dis_int(mt);
    std::vector<double> w(n), w2(n), w5(n), w4(n);
    for (int i = 0; i < n; i++) {
        double v = sqd(X + (size_t)i * d, X + (size_t)first * d, d);
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
                    w4[i] = sqd(X + (size_t)i * d, X + (size_t)cand[t] * d, d);
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

/*
 * kmeans<T>::compute restated (kmeans.cpp:351-783) for Lloyd/Elkan with initialisation
 * 0 random (std::sample, :1404-1425), 1 k-means++, 2 supplied (C_init), 3 random
 * partitions (:1426-1457). empty_mode: 0 ignore, 1 error, 2 split. Returns 0 on success,
 * 900 when every run hit an empty cluster (error mode), 21 when the best run hit max_iter.
 * out_info = {n_iter, inertia}. Centres row-major k x d.
 */
int oracle_kmeans(int n, int d, const double *X, int k, int init, const double *C_init,
                  int n_init, int max_iter, double tol, long long seed, int empty_mode,
                  double *C_out, int *labels_out, double *out_info) {
    if (n_init > 1 && init == 2)
        n_init = 1;
    std::mt19937_64 mt;
    if (init != 2)
        mt.seed((unsigned long long)seed);
    std::vector<double> C((size_t)k * d);
    std::vector<int> lab(n);
    double best = std::numeric_limits<double>::infinity();
    bool valid = false, warn = false;
    for (int run = 0; run < n_init; run++) {
        if (init == 2) {
            std::copy(C_init, C_init + (size_t)k * d, C.begin());
        } else if (init == 0) {
            std::vector<int> all(n), pick(k);
            std::iota(all.begin(), all.end(), 0);
            std::sample(all.begin(), all.end(), pick.begin(), k, mt);
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
        } else {
            kmeanspp(n, d, k, X, C.data(), mt);
        }
        KMOut o = perform(n, d, k, X, C.data(), lab.data(), max_iter, tol, empty_mode);
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
        }
    }
    if (!valid)
        return 900;
    return warn ? 21 : 0;
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

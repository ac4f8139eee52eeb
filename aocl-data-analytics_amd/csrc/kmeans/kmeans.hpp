// k-means handle (the role of da_kmeans::kmeans<T>, source/core/clustering/kmeans/
// kmeans.hpp:51-312 of the reference) with the data matrix resident in GPU memory.
#pragma once
#include "../core/handle.hpp"
#include <random>
#include <vector>

namespace dagpu {

template <typename T> class KMeansHandle : public AlgHandle {
  public:
    KMeansHandle(ErrorStack &err, GpuContext &gpu);
    da_status set_data(da_int n_samples, da_int n_features, const T *A, da_int lda);
    // Data already in device memory (row-major n_samples x n_features, copied): for callers
    // inside the library that keep their rows in HBM.
    da_status set_data_device(da_int n_samples, da_int n_features, const T *dX);
    // Results of the best run, row-major centres k x d and labels (after compute).
    const std::vector<double> &centres() const { return best_centres_; }
    const std::vector<da_int> &labels() const { return best_labels_; }
    da_status set_init_centres(const T *C, da_int ldc);
    da_status compute();
    da_status transform(da_int m_samples, da_int m_features, const T *X, da_int ldx,
                        T *X_transform, da_int ldx_transform);
    da_status predict(da_int k_samples, da_int k_features, const T *Y, da_int ldy,
                      da_int *Y_labels);
    da_status get_result_typed(da_result query, da_int *dim, T *result);
    da_status get_result(da_result query, da_int *dim, double *result) override;
    da_status get_result(da_result query, da_int *dim, float *result) override;
    da_status get_result(da_result query, da_int *dim, da_int *result) override;
    void refresh() override;
    da_status serialize(ModelBuffer &b) override;
    da_status save_check() override;
    da_status restore() override;

  private:
    struct RunResult {
        da_int n_iter = 0;
        int converged = 0;
        bool empty_found = false;
        double inertia = 0;
    };
    da_status upload_data(const T *src, da_int ld, DeviceBuffer &dst);
    da_status finish_set_data();
    da_status initialise_centres(da_int method);
    da_status kmeans_plusplus();
    da_status afk_mc2();
    // One run from the centres in Ccur_: Lloyd's iteration, or Elkan's when elkan_ is set.
    template <typename U>
    da_status run_lloyd(const U *dX, double tol, da_int max_iter, RunResult &res);
    template <typename U>
    da_status lloyd_step(const U *dX, const double *Cin, double *Cout, int *lab_new,
                         const int *lab_old, bool update, int &changed, double &normc2,
                         double &shift2);
    template <typename U>
    da_status elkan_step(const U *dX, const double *Cin, double *Cout, int *lab_new,
                         const int *lab_old, bool update, int &changed, double &normc2,
                         double &shift2);
    // Centre update from the labels lab_new (exact sums, then rounding to the lower
    // precision in the low-precision phase) and the Frobenius norms of the move.
    template <typename U>
    da_status update_centres(const U *dX, const int *lab_new, const double *Cin, double *Cout);
    // Reads back the changed-label count, the re-check count and the norms of the move.
    da_status read_step(int &changed, double &normc2, double &shift2);
    template <typename U> da_status centres_from_labels(const U *dX, const int *labels, double *C);
    template <typename U>
    da_status centre_sums(const U *dX, const int *labels, double *C, const int *esc,
                          const double *w, int sph);
    template <typename U>
    da_status handle_empty_clusters(const U *dX, double *C, int *labels, bool &split,
                                    bool &empty_error);
    template <typename U> da_status prepare_scales_for(const U *dX);
    // Fixed-point exponents of the sums for column maxima amax (see kmeans.hip) into dst.
    da_status upload_scales(const std::vector<double> &amax, DeviceBuffer &dst);
    // Spherical k-means with normalised data: 1/|x_i| per sample (else null).
    const double *winv() const { return sph_ && sph_norm_ ? inv_.as<double>() : nullptr; }
    da_status upload_rows_double(const T *src, da_int rows, da_int ld, da_order ord,
                                 DeviceBuffer &dst);

    bool initdone_ = false, centres_supplied_ = false;
    da_int n_samples_ = 0, n_features_ = 0, n_clusters_ = 1;
    da_order data_order_ = column_major;
    const T *A_usr_ = nullptr;
    da_int lda_usr_ = 0;
    const T *C_usr_ = nullptr;
    da_int ldc_usr_ = 0;
    std::vector<T> C_rowmajor_copy_;

    std::mt19937_64 mt_;
    // Device state.
    DeviceBuffer X_, Xlp_, escale_, Ca_, Cb_, lab_a_, lab_b_, counts_, part_, pcnt_, mind_,
        small_, w_, cp_, tmp_, idx_, Cbest_, q_, q_stage_, out_;
    // Spherical k-means (the "distance" option "cosine"): sph_ set by compute, sph_norm_ when
    // the data are normalised; inv_ the inverse sample norms; escale_unit_ the exponents of
    // sums of unit rows; qinv_ the inverse norms of transformed samples.
    bool sph_ = false, sph_norm_ = false;
    DeviceBuffer inv_, escale_unit_, qinv_;
    da_int ppb_ = 0, nblocks_ = 0, kc_ = 0, dl_ = 0;
    std::vector<int> escale_host_;
    // fp32 MFMA assignment: zero-padded copy of the data (n x kp_), fp32 centres, -|c|^2/2,
    // the re-check list, the centres transposed for the re-check.
    DeviceBuffer xpad_, c32_, chalf_, list_, ct_;
    int kp_ = 0;
    bool mfma_ok_ = false;
    // bf16 cross term: bf16 copy of the data padded to kp16_ (a multiple of 32) features,
    // made at the first bf16 assignment; bf16 centres.
    DeviceBuffer xb_, c16_;
    int kp16_ = 0;
    bool bf16_ok_ = false, xb_ready_ = false;

  public:
    // Samples re-checked in double by the MFMA assignments since the last compute started.
    long long rechecked_ = 0;

  private:
    da_int empty_mode_ = 0;
    // Rounding of the centres after each update: 0 none, 1 float, 2 half (the low-precision
    // phase of mixed-precision refinement).
    int lp_mode_ = 0;
    // Elkan's bounds: upper bounds (n), lower bounds (n x k), half centre distances (k x k),
    // each centre's half distance to its nearest other centre (k), centre shifts (k).
    bool elkan_ = false;
    DeviceBuffer eu_, el_, ecc_, enxt_, eshift_;
    // AFK-MC2: proposal distribution (n) and the chain's draws.
    DeviceBuffer aq_, au_;
    double *Ccur_ = nullptr, *Cprev_ = nullptr;
    int *lab_cur_ = nullptr, *lab_prev_ = nullptr;

    // Results.
    std::vector<double> best_centres_; // row-major k x d
    std::vector<da_int> best_labels_;
    double best_inertia_ = 0;
    da_int best_n_iter_ = 0, best_lp_n_iter_ = 0;
};

} // namespace dagpu

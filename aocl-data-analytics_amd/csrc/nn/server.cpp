// k-NN serving loop (aoclda_amd.h, da_nn_server_*): continuous batching of queued requests
// over a training set kept resident in GPU memory by a kNN handle.
//
// A worker thread owns two slots, each with a HIP stream, device query/result buffers and
// pinned host staging. It coalesces the queued requests (in arrival order, slicing a request
// that is larger than a batch) into a batch bounded by max_batch_queries and by the device
// bytes the slot may use, launches it on a free slot, and while that batch runs prepares
// and launches the next one on the other slot; a slot is finished (fallback of uncertified
// queries, results copied back, requests completed) before it is reused. Because every
// query's neighbours are computed exactly and independently of the other queries of its
// batch, a request's results do not depend on what it was batched with.
#include "../core/gpu.hpp"
#include "../kernels/launch.hpp"
#include "aoclda_amd.h"
#include "knn.hpp"
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

using namespace dagpu;

namespace {

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Request {
    int id = 0;
    int nq = 0, k = 0;
    std::vector<char> X; // row-major nq x d, element type of the server
    std::vector<da_int> ind;
    std::vector<char> dist;
    int next = 0;      // first query not yet batched
    int remaining = 0; // queries not yet completed
    Clock::time_point t_submit, t_done;
    bool done = false;
    bool claimed = false; // a caller is waiting for (or has taken) the result
    da_status status = da_status_success;
};

struct Segment {
    Request *req;
    int offset, count;
};

class ServerBase {
  public:
    virtual ~ServerBase() = default;
    virtual da_status submit(da_int nq, da_int d, const void *X, da_int ldx, da_int k,
                             da_int *id, int elem) = 0;
    virtual da_status wait(da_int id, da_int *ind, void *dist, double *lat, int elem) = 0;
    virtual da_status stats(double *out, da_int *n) = 0;
};

template <typename T> class Server : public ServerBase {
  public:
    Server(NNHandle<T> *nn, int max_batch, size_t max_bytes)
        : nn_(nn), max_batch_(max_batch), max_bytes_(max_bytes), t0_(Clock::now()) {}

    da_status start() {
        ErrorStack *err = nn_->errors();
        params_ = nn_->params();
        da_status s = nn_->prepare_search(params_);
        if (s != da_status_success)
            return s;
        d_ = nn_->n_features();
        // Device bytes per batched query: the query, its padded operand rows (up to three
        // copies of the row for the split forms), the per-split candidates (at most
        // knn_merge_max_candidates() (key, index) pairs, 9 bytes each with the split
        // thresholds), the fallback slots and the results.
        per_query_bytes_ = (size_t)d_ * sizeof(T) * 4 + (size_t)knn_merge_max_candidates() * 9 +
                           64 + (size_t)kmax_ * 16;
        const size_t cap = std::max<size_t>(1, max_bytes_ / 2 / per_query_bytes_);
        batch_cap_ = (int)std::min<size_t>(max_batch_, cap);
        for (auto &sl : slots_) {
            s = sl.gpu.init(err);
            if (s != da_status_success)
                return s;
            sl.ctx.gpu = &sl.gpu;
            DA_HIP(err, sl.dq.reserve((size_t)batch_cap_ * d_ * sizeof(T)));
            DA_HIP(err, hipHostMalloc((void **)&sl.hq, (size_t)batch_cap_ * d_ * sizeof(T)));
        }
        // Both slots are driven by the one worker thread: they share one set of tuned split
        // counts, so each (list length, batch size class) is tuned once, not once per slot.
        slots_[1].ctx.tuning = &slots_[0].ctx;
        worker_ = std::thread([this] { run(); });
        return da_status_success;
    }

    ~Server() override {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        if (worker_.joinable())
            worker_.join();
        for (auto &sl : slots_) {
            if (sl.hq)
                (void)hipHostFree(sl.hq);
            if (sl.hind)
                (void)hipHostFree(sl.hind);
            if (sl.hdist)
                (void)hipHostFree(sl.hdist);
        }
    }

    da_status submit(da_int nq, da_int d, const void *X, da_int ldx, da_int k, da_int *id,
                     int elem) override {
        if (elem != (int)sizeof(T))
            return da_status_wrong_type;
        if (!X || !id)
            return da_status_invalid_pointer;
        if (nq < 1 || d != d_)
            return da_status_invalid_array_dimension;
        if (ldx < d)
            return da_status_invalid_leading_dimension;
        if (k < 1 || k > nn_->n_samples() || k > kmax_)
            return da_status_invalid_input;
        auto r = std::make_unique<Request>();
        r->nq = nq;
        r->k = k;
        r->X.resize((size_t)nq * d * sizeof(T));
        const T *src = static_cast<const T *>(X);
        T *dst = reinterpret_cast<T *>(r->X.data());
        for (da_int i = 0; i < nq; i++)
            std::memcpy(dst + (size_t)i * d, src + (size_t)i * ldx, d * sizeof(T));
        r->ind.resize((size_t)nq * k);
        r->dist.resize((size_t)nq * k * sizeof(T));
        r->remaining = nq;
        r->t_submit = Clock::now();
        {
            std::lock_guard<std::mutex> g(mu_);
            r->id = next_id_++;
            *id = r->id;
            queue_.push_back(r.get());
            all_[r->id] = std::move(r);
        }
        cv_.notify_all();
        return da_status_success;
    }

    da_status wait(da_int id, da_int *ind, void *dist, double *lat, int elem) override {
        if (elem != (int)sizeof(T))
            return da_status_wrong_type;
        std::unique_ptr<Request> r;
        {
            std::unique_lock<std::mutex> g(mu_);
            auto it = all_.find(id);
            // A result is handed to one waiter only: a second wait on the same request
            // (also one that started while the first was still waiting) is refused.
            if (it == all_.end() || it->second->claimed)
                return da_status_invalid_input;
            Request *rp = it->second.get();
            rp->claimed = true;
            done_cv_.wait(g, [&] { return rp->done; });
            // std::map iterators stay valid across the insertions of concurrent submits, and
            // only the claiming waiter erases this entry.
            r = std::move(it->second);
            all_.erase(it);
        }
        if (r->status != da_status_success)
            return r->status;
        if (ind)
            std::memcpy(ind, r->ind.data(), r->ind.size() * sizeof(da_int));
        if (dist)
            std::memcpy(dist, r->dist.data(), r->dist.size());
        if (lat)
            *lat = ms_since(r->t_submit, r->t_done);
        return da_status_success;
    }

    da_status stats(double *out, da_int *n) override {
        if (!out || !n)
            return da_status_invalid_pointer;
        if (*n < 8) {
            *n = 8;
            return da_status_invalid_array_dimension;
        }
        std::lock_guard<std::mutex> g(mu_);
        std::vector<double> lat = latencies_;
        std::sort(lat.begin(), lat.end());
        auto pct = [&](double p) {
            if (lat.empty())
                return 0.0;
            const size_t i = std::min(lat.size() - 1, (size_t)(p * (lat.size() - 1) + 0.5));
            return lat[i];
        };
        out[0] = (double)completed_;
        out[1] = (double)queries_done_;
        out[2] = (double)batches_;
        out[3] = pct(0.50);
        out[4] = pct(0.99);
        out[5] = batches_ ? (double)queries_batched_ / batches_ : 0.0;
        out[6] = ms_since(t0_, Clock::now()) / 1e3;
        out[7] = (double)fallback_;
        *n = 8;
        return da_status_success;
    }

    static constexpr int kmax_ = 128;

  private:
    struct Slot {
        GpuContext gpu;
        SearchContext ctx;
        DeviceBuffer dq, dind, ddist;
        T *hq = nullptr;
        da_int *hind = nullptr;
        T *hdist = nullptr;
        size_t hres_cap = 0;
        std::vector<Segment> segs;
        int nq = 0, k = 0;
        bool busy = false, fused = false;
        da_status launch_status = da_status_success;
    };

    // Takes queued queries up to the batch capacity: the oldest request first, then, in
    // arrival order, every request whose k fits the candidate-list length of the oldest one
    // (a batch runs at the cost of its largest list, so k = 1 requests do not ride in a
    // k = 100 sweep unless the k = 100 request is the oldest). Requests are sliced at the
    // capacity; the rest of a slice stays queued in place.
    bool form_batch(std::vector<Segment> &segs, int &nq, int &k) {
        segs.clear();
        nq = 0;
        k = 1;
        if (queue_.empty())
            return false;
        const int kc = params_.tile_kc(queue_.front()->k);
        for (auto it = queue_.begin(); it != queue_.end() && nq < batch_cap_;) {
            Request *r = *it;
            if (params_.tile_kc(r->k) > kc) {
                ++it;
                continue;
            }
            const int take = std::min(r->nq - r->next, batch_cap_ - nq);
            segs.push_back({r, r->next, take});
            r->next += take;
            nq += take;
            k = std::max(k, r->k);
            if (r->next == r->nq)
                it = queue_.erase(it);
            else
                ++it;
        }
        return nq > 0;
    }

    da_status launch(Slot &sl) {
        ErrorStack *err = &werr_;
        hipStream_t st = sl.gpu.stream();
        size_t off = 0;
        for (auto &sg : sl.segs) {
            const size_t bytes = (size_t)sg.count * d_ * sizeof(T);
            std::memcpy(reinterpret_cast<char *>(sl.hq) + off,
                        sg.req->X.data() + (size_t)sg.offset * d_ * sizeof(T), bytes);
            off += bytes;
        }
        DA_HIP(err, hipMemcpyAsync(sl.dq.ptr(), sl.hq, off, hipMemcpyHostToDevice, st));
        const size_t nres = (size_t)sl.nq * sl.k;
        DA_HIP(err, sl.dind.reserve(nres * sizeof(da_int)));
        DA_HIP(err, sl.ddist.reserve(nres * sizeof(T)));
        if (sl.hres_cap < nres) {
            if (sl.hind)
                (void)hipHostFree(sl.hind);
            if (sl.hdist)
                (void)hipHostFree(sl.hdist);
            DA_HIP(err, hipHostMalloc((void **)&sl.hind, nres * sizeof(da_int)));
            DA_HIP(err, hipHostMalloc((void **)&sl.hdist, nres * sizeof(T)));
            sl.hres_cap = nres;
        }
        return nn_->search_launch(sl.ctx, params_, sl.dq.template as<T>(), sl.nq, sl.k, true,
                                  sl.dind.template as<da_int>(), sl.ddist.template as<T>(), sl.fused);
    }

    da_status finish(Slot &sl) {
        ErrorStack *err = &werr_;
        hipStream_t st = sl.gpu.stream();
        da_status s = sl.launch_status;
        if (s == da_status_success)
            s = nn_->search_finish(sl.ctx, sl.dq.template as<T>(), sl.nq, sl.k, true,
                                   sl.dind.template as<da_int>(), sl.ddist.template as<T>(),
                                   sl.fused);
        else
            (void)hipStreamSynchronize(st);
        if (s == da_status_success) {
            const size_t nres = (size_t)sl.nq * sl.k;
            DA_HIP(err, hipMemcpyAsync(sl.hind, sl.dind.ptr(), nres * sizeof(da_int),
                                       hipMemcpyDeviceToHost, st));
            DA_HIP(err, hipMemcpyAsync(sl.hdist, sl.ddist.ptr(), nres * sizeof(T),
                                       hipMemcpyDeviceToHost, st));
            DA_HIP(err, hipStreamSynchronize(st));
        }
        const auto now = Clock::now();
        std::lock_guard<std::mutex> g(mu_);
        int row = 0;
        for (auto &sg : sl.segs) {
            Request *r = sg.req;
            if (s == da_status_success) {
                T *rd = reinterpret_cast<T *>(r->dist.data());
                for (int i = 0; i < sg.count; i++) {
                    const size_t src = (size_t)(row + i) * sl.k;
                    const size_t dst = (size_t)(sg.offset + i) * r->k;
                    std::memcpy(&r->ind[dst], sl.hind + src, r->k * sizeof(da_int));
                    std::memcpy(rd + dst, sl.hdist + src, r->k * sizeof(T));
                }
            } else {
                r->status = s;
            }
            row += sg.count;
            r->remaining -= sg.count;
            if (r->remaining == 0) {
                r->t_done = now;
                r->done = true;
                completed_++;
                queries_done_ += r->nq;
                latencies_.push_back(ms_since(r->t_submit, now));
            }
        }
        batches_++;
        queries_batched_ += sl.nq;
        fallback_ += sl.ctx.last_fallback;
        sl.busy = false;
        done_cv_.notify_all();
        return s;
    }

    void run() {
        int cur = 0;
        for (;;) {
            Slot &sl = slots_[cur];
            Slot &other = slots_[cur ^ 1];
            if (sl.busy)
                finish(sl);
            bool have;
            {
                std::unique_lock<std::mutex> g(mu_);
                // With the other slot in flight, do not block: finish it if there is
                // nothing new to launch.
                if (queue_.empty() && !stop_ && !other.busy)
                    cv_.wait(g, [&] { return !queue_.empty() || stop_; });
                have = form_batch(sl.segs, sl.nq, sl.k);
                if (!have && stop_ && !other.busy)
                    return;
            }
            if (have) {
                sl.busy = true;
                // A failed launch is reported to the batch's requests by finish().
                sl.launch_status = launch(sl);
            } else if (other.busy) {
                finish(other);
                continue;
            }
            cur ^= 1;
        }
    }

    NNHandle<T> *nn_;
    int max_batch_;
    size_t max_bytes_;
    int batch_cap_ = 1;
    int d_ = 0;
    size_t per_query_bytes_ = 0;
    SearchParams params_;
    Slot slots_[2];
    std::thread worker_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    std::deque<Request *> queue_;
    std::map<int, std::unique_ptr<Request>> all_;
    // Messages of the serving thread's own device calls (the handle's stack belongs to the
    // caller's thread).
    ErrorStack werr_;
    int next_id_ = 1;
    bool stop_ = false;
    long long completed_ = 0, queries_done_ = 0, batches_ = 0, queries_batched_ = 0,
              fallback_ = 0;
    std::vector<double> latencies_;
    Clock::time_point t0_;
};

} // namespace

struct _da_nn_server {
    std::unique_ptr<ServerBase> impl;
};

extern "C" {

da_status da_nn_server_create(da_handle handle, da_int max_batch_queries,
                              size_t max_device_bytes, da_nn_server *server) {
    if (!handle || !server)
        return da_status_invalid_pointer;
    if (max_batch_queries < 1 || max_device_bytes < 1)
        return da_status_invalid_input;
    handle->err.clear();
    auto out = std::make_unique<_da_nn_server>();
    da_status s = da_status_invalid_handle_type;
    if (auto *nn = dynamic_cast<NNHandle<float> *>(handle->alg.get())) {
        if (!nn->trained())
            return da_status_no_data;
        auto sv = std::make_unique<Server<float>>(nn, max_batch_queries, max_device_bytes);
        s = sv->start();
        out->impl = std::move(sv);
    } else if (auto *nn = dynamic_cast<NNHandle<double> *>(handle->alg.get())) {
        if (!nn->trained())
            return da_status_no_data;
        auto sv = std::make_unique<Server<double>>(nn, max_batch_queries, max_device_bytes);
        s = sv->start();
        out->impl = std::move(sv);
    }
    if (s != da_status_success)
        return s;
    *server = out.release();
    return da_status_success;
}

da_status da_nn_server_submit_s(da_nn_server server, da_int n_queries, da_int n_features,
                                const float *X, da_int ldx, da_int k, da_int *request_id) {
    if (!server)
        return da_status_invalid_pointer;
    return server->impl->submit(n_queries, n_features, X, ldx, k, request_id, sizeof(float));
}
da_status da_nn_server_submit_d(da_nn_server server, da_int n_queries, da_int n_features,
                                const double *X, da_int ldx, da_int k, da_int *request_id) {
    if (!server)
        return da_status_invalid_pointer;
    return server->impl->submit(n_queries, n_features, X, ldx, k, request_id, sizeof(double));
}
da_status da_nn_server_wait_s(da_nn_server server, da_int request_id, da_int *n_ind,
                              float *n_dist, double *latency_ms) {
    if (!server)
        return da_status_invalid_pointer;
    return server->impl->wait(request_id, n_ind, n_dist, latency_ms, sizeof(float));
}
da_status da_nn_server_wait_d(da_nn_server server, da_int request_id, da_int *n_ind,
                              double *n_dist, double *latency_ms) {
    if (!server)
        return da_status_invalid_pointer;
    return server->impl->wait(request_id, n_ind, n_dist, latency_ms, sizeof(double));
}
da_status da_nn_server_stats(da_nn_server server, double *stats, da_int *n_stats) {
    if (!server)
        return da_status_invalid_pointer;
    return server->impl->stats(stats, n_stats);
}
void da_nn_server_destroy(da_nn_server *server) {
    if (server && *server) {
        delete *server;
        *server = nullptr;
    }
}

} // extern "C"

// pqivf.cpp -- the PQIVF index behind the C-ABI: host bookkeeping with the
// reference's semantics (pqivf/index.go), GPU training and GPU search.
//
// Host side (authoritative state, same rules as the reference):
//   lists[c]      entries of coarse list c in storage order (ids, vectors,
//                 codes); `exists` mirrors the presence of the map key
//                 invertedLists[c] (pqivf/index.go:43)
//   id_to_cluster id -> list (idToCluster)
//   pending       vectors added before training (pendingVectors)
// Device side (rebuilt from the host state when it changes):
//   list_start/list_len per list, ids, PQ codes (uint8, or uint16 when a
//   codebook has more than 256 rows), a has-codes flag, the raw vectors,
//   the coarse centroids and the codebooks.
#include "pqivf.h"

#include <algorithm>
#include <memory>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>

#include "batch.h"
#include "ivf.h"
#include "kmeans.h"
#include "rng.h"
#include "scan_fused.h"
#include "scan_list.h"

namespace hann {

// ------------------------------------------------------------ k-means --
void assign_points(const float *data, int64_t n, const float *cents, int k, int d, int precision, const int32_t *hint,
                   int32_t *out, hipStream_t st) {
    if (precision == HANN_PREC_BF16 && d >= 32) {
        Context &c = global_ctx();
        std::lock_guard<std::mutex> lk(c.mu);
        assign_bf16(c, data, n, cents, k, d, hint, out, st);
        return;
    }
    launch_assign(data, n, d, cents, k, d, out, st);
}

int64_t run_kmeans(const float *data, int64_t n, int64_t row_stride, int dsub, int k, int iters, int groups,
                   std::vector<uint64_t> &states, float *cents, KMeansScratch &sc, hipStream_t st, int precision,
                   int init_mode, int32_t *assign_out) {
    const int64_t keff = n < k ? n : k;
    if (keff <= 0) return 0;
    if (init_mode == 1) {
        // k-means++ (build extension): per group, a uniform first row and
        // keff - 1 uniform draws for the D^2 sampling, from the group's
        // stream; the groups' seedings run side by side
        std::vector<double> u((size_t)groups * (keff - 1) + 1);
        std::vector<int64_t> first((size_t)groups);
        for (int g = 0; g < groups; ++g) {
            first[g] = (int64_t)rng_bounded(states[g], (uint64_t)n);
            for (int64_t j = 0; j + 1 < keff; ++j)
                u[(size_t)g * (keff - 1) + j] = (double)(rng_next(states[g]) >> 11) * (1.0 / 9007199254740992.0);
        }
        DevBuf<double> du(u.size());
        HANN_HIP(hipMemcpyAsync(du.get(), u.data(), u.size() * sizeof(double), hipMemcpyHostToDevice, st));
        launch_kmeanspp(data, n, row_stride, dsub, (int)keff, groups, first.data(), du.get(), cents, st);
    } else {
        // seeds: the first keff entries of a random permutation (pqivf/index.go:804-810)
        std::vector<int64_t> init((size_t)groups * keff);
        std::vector<int64_t> perm;
        for (int g = 0; g < groups; ++g) {
            rng_perm(states[g], n, perm);
            std::copy(perm.begin(), perm.begin() + keff, init.begin() + (size_t)g * keff);
        }
        DevBuf<int64_t> dinit(init.size());
        HANN_HIP(hipMemcpyAsync(dinit.get(), init.data(), init.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
        launch_gather_group(data, row_stride, dsub, (int)keff, groups, dinit.get(), cents, st);
        HANN_HIP(hipStreamSynchronize(st));
    }
    sc.reserve(n, (int)keff, groups);
    DevBuf<int32_t> own_assign(assign_out ? 0 : (size_t)groups * n);
    struct {
        int32_t *p;
        int32_t *get() const { return p; }
    } assign{assign_out ? assign_out : own_assign.get()};
    std::vector<int64_t> offs((size_t)groups * (keff + 1));
    std::vector<int64_t> reseed((size_t)groups * keff);
    const bool bf16 = precision == HANN_PREC_BF16 && groups == 1 && dsub >= 32 && row_stride == dsub;
    for (int it = 0; it < iters; ++it) {
        if (bf16) {
            // the previous assignment's centroid bounds every point's nearest
            assign_points(data, n, cents, (int)keff, dsub, HANN_PREC_BF16, it > 0 ? assign.get() : nullptr,
                          assign.get(), st);
        } else if (dsub < 32) {
            launch_assign_groups(data, n, row_stride, cents, keff * dsub, nullptr, (int)keff, dsub, groups,
                                 assign.get(), n, st);
        } else {
            for (int g = 0; g < groups; ++g)
                launch_assign(data + (int64_t)g * dsub, n, row_stride, cents + (int64_t)g * keff * dsub, (int)keff,
                              dsub, assign.get() + (int64_t)g * n, st);
        }
        launch_group_members(assign.get(), n, n, (int)keff, groups, sc, st);
        HANN_HIP(hipMemcpyAsync(offs.data(), sc.offs.get(), offs.size() * sizeof(int64_t), hipMemcpyDeviceToHost, st));
        HANN_HIP(hipStreamSynchronize(st));
        // empty clusters draw a replacement point, in ascending cluster order
        // (pqivf/index.go:829-837)
        for (int g = 0; g < groups; ++g)
            for (int64_t c = 0; c < keff; ++c) {
                const int64_t *o = offs.data() + (size_t)g * (keff + 1);
                reseed[(size_t)g * keff + c] = o[c + 1] == o[c] ? (int64_t)rng_bounded(states[g], (uint64_t)n) : -1;
            }
        HANN_HIP(hipMemcpyAsync(sc.reseed.get(), reseed.data(), reseed.size() * sizeof(int64_t),
                                hipMemcpyHostToDevice, st));
        launch_kmeans_sums(data, n, row_stride, dsub, (int)keff, groups, cents, keff * dsub, sc, st);
    }
    HANN_HIP(hipStreamSynchronize(st));
    return keff;
}

// -------------------------------------------------------------- index --
static int64_t default_num_subquantizers(int64_t dim) {
    // pqivf/index.go:150-157
    for (int64_t n : {8, 4, 2})
        if (dim % n == 0) return n;
    return 1;
}

PQIVF::PQIVF() = default;

PQIVF::~PQIVF() {
    for (auto &kv : inflight) {
        (void)hipEventSynchronize(kv.second);
        (void)hipEventDestroy(kv.second);
    }
    if (stream) (void)hipStreamDestroy(stream);
}

void PQIVF::note_inflight(hipStream_t s) {
    hipEvent_t &e = inflight[s];
    if (!e) HANN_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HANN_HIP(hipEventRecord(e, s));
}

void PQIVF::wait_inflight() {
    for (auto &kv : inflight) HANN_HIP(hipEventSynchronize(kv.second));
}

void PQIVF::ensure_stream() {
    if (!stream) {
        ensure_device();
        HANN_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        d_fb_counter.resize(1);
        HANN_HIP(hipMemsetAsync(d_fb_counter.get(), 0, sizeof(unsigned long long), stream));
    }
}

void PQIVF::validate_vector(size_t len) const {
    // pqivf/index.go:388-396
    if (dim == 0) fail(HANN_ERR_INVALID_ARG, "cannot add to a zero-dimension index");
    if ((int64_t)len != dim)
        fail(HANN_ERR_INVALID_ARG, strf("vector dimension %zu does not match index dimension %lld", len, (long long)dim));
}

bool PQIVF::codes_valid(const int32_t *codes) const {
    if (!have_codebooks) return false;
    for (int64_t i = 0; i < m; ++i)
        if (codes[i] < 0 || codes[i] >= cb_len[i]) return false;
    return true;
}

int64_t PQIVF::nearest_centroid(const float *vec) {
    // nearestCentroid (pqivf/index.go:162-179) on the device
    if (kc == 0) fail(HANN_ERR_INVALID_ARG, "no coarse centroids available");
    ensure_stream();
    sync_device();
    scratch_f.resize(dim);
    HANN_HIP(hipMemcpyAsync(scratch_f.get(), vec, dim * sizeof(float), hipMemcpyHostToDevice, stream));
    int32_t *out = mws.get<int32_t>(19, 1);
    launch_assign(scratch_f.get(), 1, dim, d_coarse.get(), (int)kc, (int)dim, out, stream);
    int32_t c = 0;
    HANN_HIP(hipMemcpyAsync(&c, out, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
    HANN_HIP(hipStreamSynchronize(stream));
    return c;
}

void PQIVF::encode_rows(const float *vecs, int64_t n, const int32_t *clusters, int32_t *codes_out) {
    // encodeWithCodebooks (pqivf/index.go:641-666) for n rows on the device
    if (n == 0) return;
    ensure_stream();
    sync_device();
    float *dv = mws.get<float>(0, (size_t)n * dim), *resid = mws.get<float>(2, (size_t)n * dim);
    int32_t *dc = mws.get<int32_t>(1, (size_t)n), *dcodes = mws.get<int32_t>(3, (size_t)n * m);
    HANN_HIP(hipMemcpyAsync(dv, vecs, (size_t)n * dim * sizeof(float), hipMemcpyHostToDevice, stream));
    HANN_HIP(hipMemcpyAsync(dc, clusters, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, stream));
    launch_encode(dv, n, (int)dim, d_coarse.get(), dc, (int)m, d_cb.get(), (int)pmax, d_cb_len.get(), dcodes, resid,
                  stream);
    HANN_HIP(hipMemcpyAsync(codes_out, dcodes, (size_t)n * m * sizeof(int32_t), hipMemcpyDeviceToHost, stream));
    HANN_HIP(hipStreamSynchronize(stream));
}

void PQIVF::assign_locked(const std::vector<int64_t> &ids, const std::vector<const float *> &vecs) {
    // assignLocked (pqivf/index.go:245-260) for a run of new entries, batched:
    // nearest centroid and codes on the GPU, appended in the given order.
    const int64_t n = (int64_t)ids.size();
    if (n == 0) return;
    if (kc == 0) fail(HANN_ERR_INVALID_ARG, "no coarse centroids available");
    ensure_stream();
    sync_device();
    std::vector<float> flat((size_t)n * dim);
    for (int64_t i = 0; i < n; ++i) std::memcpy(&flat[(size_t)i * dim], vecs[i], dim * sizeof(float));
    float *dv = mws.get<float>(0, flat.size());
    int32_t *dc = mws.get<int32_t>(1, (size_t)n);
    HANN_HIP(hipMemcpyAsync(dv, flat.data(), flat.size() * sizeof(float), hipMemcpyHostToDevice, stream));
    launch_assign(dv, n, dim, d_coarse.get(), (int)kc, (int)dim, dc, stream);
    std::vector<int32_t> cl((size_t)n);
    std::vector<int32_t> codes;
    bool can_encode = have_codebooks;
    for (int64_t i = 0; i < m && can_encode; ++i)
        if (cb_len[i] == 0) can_encode = false;   // "failed to encode sub-vector": stored without codes
    if (can_encode) {
        float *resid = mws.get<float>(2, flat.size());
        int32_t *dcodes = mws.get<int32_t>(3, (size_t)n * m);
        launch_encode(dv, n, (int)dim, d_coarse.get(), dc, (int)m, d_cb.get(), (int)pmax, d_cb_len.get(), dcodes,
                      resid, stream);
        codes.resize((size_t)n * m);
        HANN_HIP(hipMemcpyAsync(codes.data(), dcodes, codes.size() * sizeof(int32_t), hipMemcpyDeviceToHost, stream));
    }
    HANN_HIP(hipMemcpyAsync(cl.data(), dc, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost, stream));
    HANN_HIP(hipStreamSynchronize(stream));
    for (int64_t i = 0; i < n; ++i) {
        touch_append(cl[i]);
        ListData &L = lists[cl[i]];
        L.exists = true;
        L.ids.push_back(ids[i]);
        L.vecs.insert(L.vecs.end(), vecs[i], vecs[i] + dim);
        if (can_encode) {
            L.codes.insert(L.codes.end(), codes.begin() + (size_t)i * m, codes.begin() + (size_t)(i + 1) * m);
            L.has_codes.push_back(1);
        } else {
            L.codes.insert(L.codes.end(), (size_t)m, -1);
            L.has_codes.push_back(0);
        }
        id_to_cluster[ids[i]] = cl[i];
    }
}

void PQIVF::add_locked(int64_t id, const float *vec, size_t len) {
    // addLocked (pqivf/index.go:221-237)
    validate_vector(len);
    if (id_to_cluster.count(id)) fail(HANN_ERR_DUPLICATE_ID, strf("id %lld already exists in a cluster", (long long)id));
    if (pending.count(id)) fail(HANN_ERR_DUPLICATE_ID, strf("id %lld already exists in pending vectors", (long long)id));
    if (trained) {
        assign_locked({id}, {vec});
        return;
    }
    pending.put(id, vec);
}

void PQIVF::bulk_add(const int64_t *ids, const float *flat, const size_t *lens, size_t n) {
    // BulkAdd (pqivf/index.go:266-297): ascending id order, not atomic. On a
    // trained index the GPU assigns each run of valid vectors in one batch;
    // the first invalid id stops the call after the run before it is in.
    std::vector<size_t> off(n + 1, 0);
    for (size_t i = 0; i < n; ++i) off[i + 1] = off[i] + (lens ? lens[i] : (size_t)dim);
    std::vector<size_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return ids[a] < ids[b]; });
    std::vector<int64_t> run_ids;
    std::vector<const float *> run_vecs;
    std::unordered_map<int64_t, int> in_run;
    auto flush = [&] {
        if (!run_ids.empty()) assign_locked(run_ids, run_vecs);
        run_ids.clear();
        run_vecs.clear();
        in_run.clear();
    };
    for (size_t oi = 0; oi < n; ++oi) {
        const size_t i = order[oi];
        const int64_t id = ids[i];
        const size_t len = off[i + 1] - off[i];
        const float *vec = flat + off[i];
        if ((int64_t)len != dim) {
            flush();
            fail(HANN_ERR_INVALID_ARG, strf("vector dimension %zu does not match index dimension %lld for id %lld", len,
                                            (long long)dim, (long long)id));
        }
        if (id_to_cluster.count(id) || in_run.count(id)) {
            flush();
            fail(HANN_ERR_DUPLICATE_ID, strf("id %lld already exists in a cluster", (long long)id));
        }
        if (pending.count(id)) {
            flush();
            fail(HANN_ERR_DUPLICATE_ID, strf("id %lld already exists in pending vectors", (long long)id));
        }
        if (trained) {
            run_ids.push_back(id);
            run_vecs.push_back(vec);
            in_run[id] = 1;
        } else {
            pending.put(id, vec);
        }
    }
    flush();
}

void PQIVF::remove_from_list(int64_t id, int64_t cluster) {
    touch_delete(cluster, id);
    ListData &L = lists[cluster];
    size_t w = 0;
    const size_t len = L.ids.size();
    for (size_t r = 0; r < len; ++r) {
        if (L.ids[r] == id) continue;
        if (w != r) {
            L.ids[w] = L.ids[r];
            std::memcpy(&L.vecs[w * dim], &L.vecs[r * dim], dim * sizeof(float));
            std::memcpy(&L.codes[w * m], &L.codes[r * m], m * sizeof(int32_t));
            L.has_codes[w] = L.has_codes[r];
        }
        ++w;
    }
    L.ids.resize(w);
    L.vecs.resize(w * dim);
    L.codes.resize(w * m);
    L.has_codes.resize(w);
}

void PQIVF::delete_locked(int64_t id) {
    // deleteLocked (pqivf/index.go:313-345): order-preserving removal
    if (pending.erase(id)) return;
    auto it = id_to_cluster.find(id);
    if (it == id_to_cluster.end()) fail(HANN_ERR_NOT_FOUND, strf("id %lld not found", (long long)id));
    const int64_t c = it->second;
    if (c < 0 || c >= (int64_t)lists.size() || !lists[c].exists)
        fail(HANN_ERR_NOT_FOUND, strf("inconsistent state: cluster %lld not found for id %lld", (long long)c, (long long)id));
    const size_t before = lists[c].ids.size();
    remove_from_list(id, c);
    if (lists[c].ids.size() == before)
        fail(HANN_ERR_NOT_FOUND, strf("id %lld not found in cluster %lld", (long long)id, (long long)c));
    id_to_cluster.erase(it);
}

void PQIVF::bulk_delete(const int64_t *ids, size_t n) {
    // BulkDelete (pqivf/index.go:350-384): unknown ids are skipped
    // Removing the ids one at a time (as the reference does) or per list in
    // one pass leaves the same lists: each removal keeps the order of the rest.
    std::vector<int64_t> sorted(ids, ids + n);
    std::sort(sorted.begin(), sorted.end());
    std::vector<std::pair<int64_t, int64_t>> by_list;   // (cluster, id)
    for (int64_t id : sorted) {
        if (pending.erase(id)) continue;
        auto it = id_to_cluster.find(id);
        if (it == id_to_cluster.end()) continue;
        const int64_t c = it->second;
        if (c < 0 || c >= (int64_t)lists.size() || !lists[c].exists) continue;
        by_list.emplace_back(c, id);
        id_to_cluster.erase(it);
    }
    std::sort(by_list.begin(), by_list.end());
    // runs of one list: (list, first, end) in by_list; the bookkeeping is
    // serial, the per-list compactions are independent and run on threads
    std::vector<std::array<size_t, 3>> runs;
    for (size_t i = 0; i < by_list.size();) {
        const int64_t c = by_list[i].first;
        const size_t b = i;
        for (; i < by_list.size() && by_list[i].first == c; ++i) touch_delete(c, by_list[i].second);
        runs.push_back({(size_t)c, b, i});
    }
    auto compact = [&](size_t r0, size_t r1) {
        for (size_t k = r0; k < r1; ++k) {
            ListData &L = lists[runs[k][0]];
            const auto first = by_list.begin() + (ptrdiff_t)runs[k][1], last = by_list.begin() + (ptrdiff_t)runs[k][2];
            auto dropped = [&](int64_t id) {
                auto it = std::lower_bound(first, last, std::make_pair((int64_t)runs[k][0], id));
                return it != last && it->second == id;
            };
            size_t w = 0;
            const size_t len = L.ids.size();
            for (size_t r = 0; r < len; ++r) {
                if (dropped(L.ids[r])) continue;
                if (w != r) {
                    L.ids[w] = L.ids[r];
                    std::memcpy(&L.vecs[w * dim], &L.vecs[r * dim], dim * sizeof(float));
                    std::memcpy(&L.codes[w * m], &L.codes[r * m], m * sizeof(int32_t));
                    L.has_codes[w] = L.has_codes[r];
                }
                ++w;
            }
            L.ids.resize(w);
            L.vecs.resize(w * dim);
            L.codes.resize(w * m);
            L.has_codes.resize(w);
        }
    };
    const size_t nthreads = std::min<size_t>(runs.size() / 8, std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
    if (nthreads <= 1) {
        compact(0, runs.size());
    } else {
        std::vector<std::thread> pool;
        const size_t per = (runs.size() + nthreads - 1) / nthreads;
        for (size_t t = 0; t < nthreads; ++t) {
            const size_t r0 = t * per, r1 = std::min(runs.size(), r0 + per);
            if (r0 < r1) pool.emplace_back(compact, r0, r1);
        }
        for (auto &th : pool) th.join();
    }
}

void PQIVF::update(int64_t id, const float *vec, size_t len) {
    // Update (pqivf/index.go:402-413)
    validate_vector(len);
    delete_locked(id);
    add_locked(id, vec, len);
}

void PQIVF::bulk_update(const int64_t *ids, const float *flat, const size_t *lens, size_t n) {
    // BulkUpdate (pqivf/index.go:419-444): validate all, then delete + add in
    // ascending id order
    std::vector<size_t> off(n + 1, 0);
    for (size_t i = 0; i < n; ++i) off[i + 1] = off[i] + (lens ? lens[i] : (size_t)dim);
    std::vector<size_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return ids[a] < ids[b]; });
    for (size_t i : order) {
        const size_t len = off[i + 1] - off[i];
        try {
            validate_vector(len);
        } catch (const Error &e) {
            fail(e.code, strf("invalid vector for id %lld: %s", (long long)ids[i], e.what()));
        }
    }
    for (size_t i : order) {
        delete_locked(ids[i]);
        add_locked(ids[i], flat + off[i], off[i + 1] - off[i]);
    }
}

void PQIVF::reserve_slack(ListData &L, int64_t len) const {
    // appends after a build fill this room before any reallocation (which
    // would copy the whole list) -- the host side of the device mirror's slack
    const int64_t cap = len + std::max<int64_t>(8, len / 8);
    L.ids.reserve(cap);
    L.vecs.reserve(cap * dim);
    L.codes.reserve(cap * m);
    L.has_codes.reserve(cap);
}

// threads for host-side copies of a train's vectors
static size_t host_threads(int64_t work_rows) {
    const size_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    return (size_t)std::max<int64_t>(1, std::min<int64_t>((int64_t)hw, work_rows / 65536));
}

void PQIVF::train() {
    // Train (pqivf/index.go:451-605)
    std::vector<int64_t> all_ids;
    std::vector<const float *> all_vecs;
    size_t total = pending.size();
    for (const ListData &L : lists) total += L.ids.size();
    all_ids.reserve(total);
    all_vecs.reserve(total);
    for (const ListData &L : lists)
        for (size_t r = 0; r < L.ids.size(); ++r) {
            all_ids.push_back(L.ids[r]);
            all_vecs.push_back(&L.vecs[r * dim]);
        }
    for (size_t i = 0; i < pending.size(); ++i) {
        all_ids.push_back(pending.ids[i]);
        all_vecs.push_back(&pending.vecs[i * dim]);
    }
    const int64_t n = (int64_t)all_ids.size();
    if (n < coarse_k)
        fail(HANN_ERR_EMPTY, strf("not enough vectors (%lld) to train coarse quantizer with %lld clusters", (long long)n,
                                  (long long)coarse_k));
    // ascending id order (ids are distinct); vectors added in id order need
    // no sort. The vectors stay where they are (the old lists and the
    // pending store, both alive until the new state is committed): src(i)
    // is the i-th vector in id order
    std::vector<int64_t> order;
    if (!std::is_sorted(all_ids.begin(), all_ids.end())) {
        std::vector<std::pair<int64_t, int64_t>> key((size_t)n);
        for (int64_t i = 0; i < n; ++i) key[i] = {all_ids[i], i};
        std::sort(key.begin(), key.end());
        order.resize((size_t)n);
        for (int64_t i = 0; i < n; ++i) order[i] = key[i].second;
    }
    auto src = [&](int64_t i) { return all_vecs[order.empty() ? i : order[i]]; };
    std::vector<int64_t> ids((size_t)n);
    for (int64_t i = 0; i < n; ++i) ids[i] = all_ids[order.empty() ? i : order[i]];
    ensure_stream();
    const uint64_t epoch = train_epoch;
    const int d = (int)dim;
    DevBuf<float> dX((size_t)n * dim);
    {
        // rows that sit one after another in host memory go up in one copy;
        // many short runs are packed on the host first (in parallel)
        std::vector<std::pair<int64_t, int64_t>> runs;   // (first row, rows)
        for (int64_t i = 0; i < n;) {
            int64_t j = i + 1;
            while (j < n && src(j) == src(j - 1) + dim) ++j;
            runs.push_back({i, j - i});
            i = j;
            if (runs.size() > 4096) break;
        }
        if (runs.size() <= 4096) {
            for (const auto &r : runs)
                HANN_HIP(hipMemcpyAsync(dX.get() + r.first * dim, src(r.first), (size_t)r.second * dim * sizeof(float),
                                        hipMemcpyHostToDevice, stream));
        } else {
            std::vector<float> X((size_t)n * dim);
            const size_t nt = host_threads(n);
            std::vector<std::thread> pool;
            for (size_t t = 0; t < nt; ++t)
                pool.emplace_back([&, t] {
                    for (int64_t i = (int64_t)t * n / (int64_t)nt; i < (int64_t)(t + 1) * n / (int64_t)nt; ++i)
                        std::memcpy(&X[(size_t)i * dim], src(i), dim * sizeof(float));
                });
            for (auto &th : pool) th.join();
            HANN_HIP(hipMemcpyAsync(dX.get(), X.data(), X.size() * sizeof(float), hipMemcpyHostToDevice, stream));
            HANN_HIP(hipStreamSynchronize(stream));
        }
    }

    // coarse quantizer, optionally on a seeded sample (build extension)
    KMeansScratch sc;
    const float *train_src = dX.get();
    int64_t ntrain = n;
    DevBuf<float> dS;
    if (coarse_sample > 0 && coarse_sample < n) {
        uint64_t st = stream_seed(seed, epoch, kRoleCoarseSample);
        std::vector<int64_t> perm;
        rng_perm(st, n, perm);
        std::vector<int64_t> pick(perm.begin(), perm.begin() + coarse_sample);
        std::sort(pick.begin(), pick.end());
        DevBuf<int64_t> dpick(pick.size());
        HANN_HIP(hipMemcpyAsync(dpick.get(), pick.data(), pick.size() * sizeof(int64_t), hipMemcpyHostToDevice, stream));
        dS.resize((size_t)coarse_sample * dim);
        launch_gather_rows(dX.get(), d, dpick.get(), coarse_sample, dS.get(), stream);
        HANN_HIP(hipStreamSynchronize(stream));
        train_src = dS.get();
        ntrain = coarse_sample;
    }
    DevBuf<float> dC((size_t)coarse_k * dim);
    std::vector<uint64_t> st0{stream_seed(seed, epoch, 0)};
    // k-means over every vector leaves its last assignment in dA, a bound
    // for the final one (the certified BF16 assignment's hint)
    DevBuf<int32_t> dA((size_t)n);
    const bool hinted = ntrain == n && iters > 0;
    const int64_t new_kc =
        run_kmeans(train_src, ntrain, dim, d, (int)coarse_k, (int)iters, 1, st0, dC.get(), sc, stream, train_precision,
                   kmeans_init, hinted ? dA.get() : nullptr);
    dS.release();

    // final assignment of every vector; lists in ascending id order
    assign_points(dX.get(), n, dC.get(), (int)new_kc, d, train_precision, hinted ? dA.get() : nullptr, dA.get(),
                  stream);
    std::vector<int32_t> A((size_t)n);
    HANN_HIP(hipMemcpyAsync(A.data(), dA.get(), A.size() * sizeof(int32_t), hipMemcpyDeviceToHost, stream));

    // PQ training rows: residuals in cluster-ascending, list order
    // (pqivf/index.go:520-541) = the stable grouping of points by cluster
    KMeansScratch sg;
    sg.reserve(n, (int)new_kc, 1);
    launch_group_members(dA.get(), n, n, (int)new_kc, 1, sg, stream);
    int64_t npq = n;
    const int32_t *pq_order = sg.members.get();
    DevBuf<int32_t> dsub_order;
    std::vector<int32_t> full, sub;   // host staging of the sample order, alive until the stream syncs
    if (pq_sample > 0 && pq_sample < n) {
        uint64_t st = stream_seed(seed, epoch, kRolePQSample);
        std::vector<int64_t> perm;
        rng_perm(st, n, perm);
        std::vector<int64_t> pick(perm.begin(), perm.begin() + pq_sample);
        std::sort(pick.begin(), pick.end());
        full.resize((size_t)n);
        sub.resize((size_t)pq_sample);
        HANN_HIP(hipMemcpyAsync(full.data(), sg.members.get(), (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost, stream));
        HANN_HIP(hipStreamSynchronize(stream));
        for (int64_t i = 0; i < pq_sample; ++i) sub[i] = full[pick[i]];
        dsub_order.resize(sub.size());
        HANN_HIP(hipMemcpyAsync(dsub_order.get(), sub.data(), sub.size() * sizeof(int32_t), hipMemcpyHostToDevice, stream));
        pq_order = dsub_order.get();
        npq = pq_sample;
    }
    DevBuf<float> dR((size_t)npq * dim);
    launch_residuals(dX.get(), d, dC.get(), dA.get(), pq_order, npq, dR.get(), stream);
    const int s = (int)(dim / m);
    DevBuf<float> dCB((size_t)m * pq_k * s);
    std::vector<uint64_t> sts((size_t)m);
    for (int64_t i = 0; i < m; ++i) sts[i] = stream_seed(seed, epoch, 1 + (uint64_t)i);
    KMeansScratch sp;
    const int64_t pk = run_kmeans(dR.get(), npq, dim, s, (int)pq_k, (int)iters, (int)m, sts, dCB.get(), sp, stream,
                                  HANN_PREC_FP32, kmeans_init);
    dR.release();

    // encode every vector (pqivf/index.go:554-601)
    std::vector<int64_t> cbl((size_t)m, pk);
    DevBuf<int64_t> dcbl((size_t)m);
    HANN_HIP(hipMemcpyAsync(dcbl.get(), cbl.data(), cbl.size() * sizeof(int64_t), hipMemcpyHostToDevice, stream));
    DevBuf<int32_t> dcodes((size_t)n * m);
    {
        DevBuf<float> resid(s >= 32 ? (size_t)n * dim : 1);
        launch_encode(dX.get(), n, d, dC.get(), dA.get(), (int)m, dCB.get(), (int)pk, dcbl.get(), dcodes.get(),
                      resid.get(), stream);
    }
    std::vector<int32_t> codes((size_t)n * m);
    std::vector<float> C((size_t)new_kc * dim), CB((size_t)m * pk * s);
    HANN_HIP(hipMemcpyAsync(codes.data(), dcodes.get(), codes.size() * sizeof(int32_t), hipMemcpyDeviceToHost, stream));
    HANN_HIP(hipMemcpyAsync(C.data(), dC.get(), C.size() * sizeof(float), hipMemcpyDeviceToHost, stream));
    HANN_HIP(hipMemcpyAsync(CB.data(), dCB.get(), CB.size() * sizeof(float), hipMemcpyDeviceToHost, stream));
    HANN_HIP(hipStreamSynchronize(stream));

    // commit the new state
    train_epoch = epoch + 1;
    kc = new_kc;
    coarse.swap(C);
    codebooks.swap(CB);
    pmax = pk;
    cb_len = cbl;
    have_codebooks = true;
    // the new lists: each list's members in ascending id order, built by
    // threads over ranges of lists while one more fills the id map (the
    // vectors are read from their old places, freed after the swap)
    std::vector<ListData> fresh((size_t)kc);
    std::vector<int64_t> off((size_t)kc + 1, 0);
    for (int64_t i = 0; i < n; ++i) off[(size_t)A[i] + 1]++;
    for (int64_t c = 0; c < kc; ++c) off[c + 1] += off[c];
    std::vector<int64_t> members((size_t)n);
    {
        std::vector<int64_t> pos(off.begin(), off.end() - 1);
        for (int64_t i = 0; i < n; ++i) members[pos[A[i]]++] = i;
    }
    {
        const size_t nt = host_threads(n);
        std::vector<std::thread> pool;
        pool.emplace_back([&] {
            id_to_cluster.clear();
            id_to_cluster.reserve((size_t)n);
            for (int64_t i = 0; i < n; ++i) id_to_cluster[ids[i]] = A[i];
        });
        // list ranges of about n / nt members each
        std::vector<int64_t> cut{0};
        for (size_t t = 1; t < nt; ++t) {
            const int64_t want = (int64_t)t * n / (int64_t)nt;
            cut.push_back(std::lower_bound(off.begin(), off.end(), want) - off.begin());
        }
        cut.push_back(kc);
        for (size_t t = 0; t < nt; ++t)
            pool.emplace_back([&, t] {
                for (int64_t c = cut[t]; c < cut[t + 1]; ++c) {
                    ListData &L = fresh[c];
                    const int64_t len = off[c + 1] - off[c];
                    reserve_slack(L, len);
                    for (int64_t e = off[c]; e < off[c + 1]; ++e) {
                        const int64_t i = members[e];
                        L.exists = true;
                        L.ids.push_back(ids[i]);
                        const float *v = src(i);
                        L.vecs.insert(L.vecs.end(), v, v + dim);
                        L.codes.insert(L.codes.end(), &codes[(size_t)i * m], &codes[(size_t)(i + 1) * m]);
                        L.has_codes.push_back(1);
                    }
                }
            });
        for (auto &th : pool) th.join();
    }
    lists.swap(fresh);
    pending.clear();
    trained = true;
    mark_dirty();
}

// ------------------------------------------------------- device mirror --
void PQIVF::mark_dirty() {
    dirty = true;
    full_rebuild = true;
}

void PQIVF::touch_append(int64_t l) {
    dirty = true;
    if (full_rebuild) return;
    if (touch_flags[l] == 0) {
        touch_flags[l] = 1;
        append_from[l] = (int64_t)lists[l].ids.size();
        touched_lists.push_back(l);
    }
}

void PQIVF::touch_other(int64_t l) {
    dirty = true;
    if (full_rebuild) return;
    if (touch_flags[l] == 0) touched_lists.push_back(l);
    touch_flags[l] |= 4;
}

void PQIVF::touch_delete(int64_t l, int64_t id) {
    dirty = true;
    if (full_rebuild) return;
    if (touch_flags[l] == 0) touched_lists.push_back(l);
    touch_flags[l] |= 2;
    deleted_ids[l].push_back(id);
}

// Stable removal of deleted ids from lists on the device, in two steps.
// k_compact_map: job j covers list slots [start[j], start[j] + len[j]) and
// drops the ids del[doff[j] .. doff[j+1]) (sorted); its kept entries, in
// order, get rows koff[j] .. of a packed map: smap = their slot, dmap = the
// slot they move to (start[j] + rank among the kept).
__global__ void __launch_bounds__(256) k_compact_map(const int64_t *__restrict__ start, const int64_t *__restrict__ len,
                                                     const int64_t *__restrict__ doff, const int64_t *__restrict__ del,
                                                     const int64_t *__restrict__ koff, const int64_t *__restrict__ ids,
                                                     int64_t *__restrict__ smap, int64_t *__restrict__ dmap) {
    __shared__ int s_scan[256];
    const int j = blockIdx.x;
    const int64_t s0 = start[j], n = len[j];
    const int64_t *dl = del + doff[j];
    const int64_t nd = doff[j + 1] - doff[j];
    int64_t run = 0;
    for (int64_t c0 = 0; c0 < n; c0 += 256) {
        const int64_t e = c0 + threadIdx.x;
        int keep = 0;
        if (e < n) {
            const int64_t id = ids[s0 + e];
            int64_t lo = 0, hi = nd;
            while (lo < hi) {
                const int64_t mid = (lo + hi) >> 1;
                if (dl[mid] < id) lo = mid + 1; else hi = mid;
            }
            keep = !(lo < nd && dl[lo] == id);
        }
        s_scan[threadIdx.x] = keep;
        __syncthreads();
        for (int o = 1; o < 256; o <<= 1) {
            const int v = (int)threadIdx.x >= o ? s_scan[threadIdx.x - o] : 0;
            __syncthreads();
            s_scan[threadIdx.x] += v;
            __syncthreads();
        }
        if (keep) {
            const int64_t rank = run + s_scan[threadIdx.x] - 1;
            smap[koff[j] + rank] = s0 + e;
            dmap[koff[j] + rank] = s0 + rank;
        }
        run += s_scan[255];
        __syncthreads();
    }
}

// Entries rows [0, n): from slot smap[i] (or row i) to slot dmap[i] (or
// row i); vectors and codes move coalesced (a thread per element).
__global__ void k_move_entries(int64_t n, int cb, int d, const int64_t *__restrict__ smap,
                               const int64_t *__restrict__ dmap, const int64_t *__restrict__ ids,
                               const uint8_t *__restrict__ codes, const uint8_t *__restrict__ has,
                               const float *__restrict__ vecs, int64_t *__restrict__ o_ids, uint8_t *__restrict__ o_codes,
                               uint8_t *__restrict__ o_has, float *__restrict__ o_vecs) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t t = t0; t < n * d; t += stride) {
        const int64_t r = t / d;
        const int x = (int)(t - r * d);
        const int64_t src = smap ? smap[r] : r, dst = dmap ? dmap[r] : r;
        o_vecs[dst * d + x] = vecs[src * d + x];
    }
    for (int64_t t = t0; t < n * cb; t += stride) {
        const int64_t r = t / cb;
        const int x = (int)(t - r * cb);
        const int64_t src = smap ? smap[r] : r, dst = dmap ? dmap[r] : r;
        o_codes[dst * cb + x] = codes[src * cb + x];
    }
    for (int64_t r = t0; r < n; r += stride) {
        const int64_t src = smap ? smap[r] : r, dst = dmap ? dmap[r] : r;
        o_ids[dst] = ids[src];
        o_has[dst] = has[src];
    }
}

// dst[i] = slot of packed entry i: copy its id, codes, flag and vector there
__global__ void k_scatter_entries(int64_t n, int m, int code_bytes, int d, const int64_t *__restrict__ dst,
                                  const int64_t *__restrict__ ids, const uint8_t *__restrict__ codes,
                                  const uint8_t *__restrict__ has, const float *__restrict__ vecs,
                                  int64_t *__restrict__ o_ids, uint8_t *__restrict__ o_codes,
                                  uint8_t *__restrict__ o_has, float *__restrict__ o_vecs) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * d) return;
    const int64_t i = t / d;
    const int e = (int)(t % d);
    const int64_t p = dst[i];
    o_vecs[p * d + e] = vecs[t];
    if (e == 0) {
        o_ids[p] = ids[i];
        o_has[p] = has[i];
        const int cb = m * code_bytes;
        for (int b = 0; b < cb; ++b) o_codes[p * cb + b] = codes[i * cb + b];
    }
}

void PQIVF::sync_device() {
    ensure_stream();
    if (!dirty) return;
    wait_inflight();
    ++mirror_epoch;
    const int s = m > 0 ? (int)(dim / m) : 0;
    code_bytes = pmax > 256 ? 2 : 1;
    auto pack_entry = [&](const ListData &L, size_t r, int64_t pos, std::vector<int64_t> &hids,
                          std::vector<uint8_t> &hcodes, std::vector<uint8_t> &hhas, std::vector<float> &hv) {
        hids[pos] = L.ids[r];
        std::memcpy(&hv[(size_t)pos * dim], &L.vecs[r * dim], dim * sizeof(float));
        const int32_t *cr = &L.codes[r * m];
        const bool ok = L.has_codes[r] && codes_valid(cr);
        hhas[pos] = ok ? 1 : 0;
        if (!ok) any_uncoded = true;
        for (int64_t i = 0; i < m; ++i) {
            const int32_t v = ok ? cr[i] : 0;
            if (code_bytes == 1)
                hcodes[(size_t)pos * m + i] = (uint8_t)v;
            else
                reinterpret_cast<uint16_t *>(hcodes.data())[(size_t)pos * m + i] = (uint16_t)v;
        }
    };
    // incremental: upload appended tails, or whole changed lists, into the
    // lists' existing device slots
    if (!full_rebuild && dev_start.size() == lists.size()) {
        bool fits = true;
        int64_t total = 0;
        std::vector<int64_t> jobs;   // delete-only lists: compacted on the device
        for (int64_t l : touched_lists) {
            const int64_t len = (int64_t)lists[l].ids.size();
            if (len > dev_cap[l]) fits = false;
            if (touch_flags[l] == 2) {
                jobs.push_back(l);
                continue;
            }
            total += len - (touch_flags[l] == 1 ? append_from[l] : 0);
        }
        if (fits && !jobs.empty()) {
            const size_t J = jobs.size();
            std::vector<int64_t> jstart(J), jlen(J), doff(J + 1, 0), koff(J + 1, 0), newlen(J), dels;
            for (size_t j = 0; j < J; ++j) {
                const int64_t l = jobs[j];
                jstart[j] = dev_start[l];
                jlen[j] = dev_len[l];
                std::vector<int64_t> &dv = deleted_ids[l];
                std::sort(dv.begin(), dv.end());
                dels.insert(dels.end(), dv.begin(), dv.end());
                doff[j + 1] = (int64_t)dels.size();
                newlen[j] = (int64_t)lists[l].ids.size();
                koff[j + 1] = koff[j] + newlen[j];
            }
            const int64_t T = koff[J];   // kept entries over the jobs
            int64_t *d_js = mws.get<int64_t>(4, J), *d_jl = mws.get<int64_t>(5, J), *d_do = mws.get<int64_t>(6, J + 1),
                    *d_del = mws.get<int64_t>(7, dels.size() + 1), *d_ko = mws.get<int64_t>(8, J + 1),
                    *smap = mws.get<int64_t>(9, (size_t)T + 1), *dmap = mws.get<int64_t>(20, (size_t)T + 1),
                    *t_ids = mws.get<int64_t>(10, (size_t)T + 1);
            uint8_t *t_codes = mws.get<uint8_t>(11, (size_t)T * m * code_bytes + 1),
                    *t_has = mws.get<uint8_t>(12, (size_t)T + 1);
            float *t_vecs = mws.get<float>(13, (size_t)T * dim + 1);
            HANN_HIP(hipMemcpyAsync(d_js, jstart.data(), J * 8, hipMemcpyHostToDevice, stream));
            HANN_HIP(hipMemcpyAsync(d_jl, jlen.data(), J * 8, hipMemcpyHostToDevice, stream));
            HANN_HIP(hipMemcpyAsync(d_do, doff.data(), (J + 1) * 8, hipMemcpyHostToDevice, stream));
            if (!dels.empty())
                HANN_HIP(hipMemcpyAsync(d_del, dels.data(), dels.size() * 8, hipMemcpyHostToDevice, stream));
            HANN_HIP(hipMemcpyAsync(d_ko, koff.data(), (J + 1) * 8, hipMemcpyHostToDevice, stream));
            hipLaunchKernelGGL(k_compact_map, dim3((unsigned)J), dim3(256), 0, stream, d_js, d_jl, d_do, d_del, d_ko,
                               d_ids.get(), smap, dmap);
            HANN_LAUNCH_CHECK();
            if (T > 0) {
                const int cbw = (int)(m * code_bytes);
                const unsigned g = (unsigned)std::min<int64_t>(65535, ceil_div(T * dim, 256));
                // kept entries to packed rows, then back to their new slots
                hipLaunchKernelGGL(k_move_entries, dim3(g), dim3(256), 0, stream, T, cbw, (int)dim, smap,
                                   (const int64_t *)nullptr, d_ids.get(), d_codes.get(), d_has_codes.get(),
                                   d_vecs.get(), t_ids, t_codes, t_has, t_vecs);
                HANN_LAUNCH_CHECK();
                hipLaunchKernelGGL(k_move_entries, dim3(g), dim3(256), 0, stream, T, cbw, (int)dim,
                                   (const int64_t *)nullptr, dmap, t_ids, t_codes, t_has, t_vecs, d_ids.get(),
                                   d_codes.get(), d_has_codes.get(), d_vecs.get());
                HANN_LAUNCH_CHECK();
            }
            HANN_HIP(hipStreamSynchronize(stream));
            for (size_t j = 0; j < J; ++j) {
                dev_len[jobs[j]] = newlen[j];
                deleted_ids[jobs[j]].clear();
                touch_flags[jobs[j]] = 0;
            }
            HANN_HIP(hipMemcpyAsync(d_list_len.get(), dev_len.data(), dev_len.size() * 8, hipMemcpyHostToDevice,
                                    stream));
            refresh_terms(&jobs);
            HANN_HIP(hipStreamSynchronize(stream));
            std::vector<int64_t> rest;
            for (int64_t l : touched_lists)
                if (touch_flags[l] != 0) rest.push_back(l);
            touched_lists.swap(rest);
        }
        if (fits) {
            const std::vector<int64_t> changed = touched_lists;
            std::vector<int64_t> hids((size_t)total), dst((size_t)total);
            std::vector<uint8_t> hcodes((size_t)total * m * code_bytes), hhas((size_t)total);
            std::vector<float> hv((size_t)total * dim);
            int64_t pos = 0;
            for (int64_t l : touched_lists) {
                const ListData &L = lists[l];
                const int64_t from = touch_flags[l] == 1 ? append_from[l] : 0;
                for (int64_t r = from; r < (int64_t)L.ids.size(); ++r, ++pos) {
                    pack_entry(L, (size_t)r, pos, hids, hcodes, hhas, hv);
                    dst[pos] = dev_start[l] + r;
                }
                dev_len[l] = (int64_t)L.ids.size();
                touch_flags[l] = 0;
                deleted_ids[l].clear();
            }
            touched_lists.clear();
            if (total > 0) {
                int64_t *d_dst = mws.get<int64_t>(14, (size_t)total), *d_i = mws.get<int64_t>(15, (size_t)total);
                uint8_t *d_c = mws.get<uint8_t>(16, hcodes.size()), *d_h = mws.get<uint8_t>(17, (size_t)total);
                float *d_v = mws.get<float>(18, hv.size());
                HANN_HIP(hipMemcpyAsync(d_dst, dst.data(), total * 8, hipMemcpyHostToDevice, stream));
                HANN_HIP(hipMemcpyAsync(d_i, hids.data(), total * 8, hipMemcpyHostToDevice, stream));
                HANN_HIP(hipMemcpyAsync(d_c, hcodes.data(), hcodes.size(), hipMemcpyHostToDevice, stream));
                HANN_HIP(hipMemcpyAsync(d_h, hhas.data(), total, hipMemcpyHostToDevice, stream));
                HANN_HIP(hipMemcpyAsync(d_v, hv.data(), hv.size() * 4, hipMemcpyHostToDevice, stream));
                hipLaunchKernelGGL(k_scatter_entries, dim3((unsigned)ceil_div(total * dim, 256)), dim3(256), 0, stream,
                                   total, (int)m, code_bytes, (int)dim, d_dst, d_i, d_c, d_h, d_v, d_ids.get(),
                                   d_codes.get(), d_has_codes.get(), d_vecs.get());
                HANN_LAUNCH_CHECK();
                HANN_HIP(hipMemcpyAsync(d_list_len.get(), dev_len.data(), dev_len.size() * 8, hipMemcpyHostToDevice,
                                        stream));
                refresh_terms(&changed);
                HANN_HIP(hipStreamSynchronize(stream));
            }
            n_entries = 0;
            for (int64_t v : dev_len) n_entries += v;
            ++incremental_syncs;
            dirty = false;
            return;
        }
    }
    // full rebuild: every list gets slack capacity for later appends
    const size_t K = lists.size();
    dev_start.assign(K, 0);
    dev_cap.assign(K, 0);
    dev_len.assign(K, 0);
    append_from.assign(K, 0);
    touch_flags.assign(K, 0);
    deleted_ids.assign(K, {});
    touched_lists.clear();
    int64_t N = 0;
    for (size_t c = 0; c < K; ++c) {
        const int64_t len = (int64_t)lists[c].ids.size();
        dev_start[c] = N;
        dev_len[c] = len;
        dev_cap[c] = len + std::max<int64_t>(8, len / 8);
        N += dev_cap[c];
    }
    n_entries = 0;
    for (int64_t v : dev_len) n_entries += v;
    std::vector<int64_t> hids((size_t)N, -1);
    std::vector<uint8_t> hcodes((size_t)N * m * code_bytes, 0);
    std::vector<uint8_t> hhas((size_t)N, 0);
    std::vector<float> hv((size_t)N * dim, 0.f);
    any_uncoded = false;
    for (size_t c = 0; c < K; ++c) {
        const ListData &L = lists[c];
        for (size_t r = 0; r < L.ids.size(); ++r) pack_entry(L, r, dev_start[c] + (int64_t)r, hids, hcodes, hhas, hv);
    }
    std::vector<int64_t> &start = dev_start;
    std::vector<int64_t> &len = dev_len;
    full_rebuild = false;
    ++full_syncs;
    auto up = [&](auto &buf, const auto &host) {
        using T = typename std::decay_t<decltype(host)>::value_type;
        buf.resize(host.size() + 1);
        if (!host.empty())
            HANN_HIP(hipMemcpyAsync(buf.get(), host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice, stream));
    };
    up(d_list_start, start);
    up(d_list_len, len);
    up(d_ids, hids);
    up(d_codes, hcodes);
    up(d_has_codes, hhas);
    up(d_vecs, hv);
    up(d_coarse, coarse);
    up(d_cb, codebooks);
    up(d_cb_len, cb_len);
    // the FP8 list scan's codebook scale: the largest power of two that
    // keeps every codeword value within 224 (half of e4m3's largest)
    {
        float mx = 0.f;
        for (float v : codebooks) mx = std::max(mx, std::fabs(v));
        cb_fp8_scale = mx > 0.f && std::isfinite(mx) ? std::exp2(std::floor(std::log2(224.f / mx))) : 1.f;
    }
    (void)s;
    coarse_bf16.ready = false;
    d_term.resize((size_t)N + 1);
    refresh_terms(nullptr);
    HANN_HIP(hipStreamSynchronize(stream));
    dirty = false;
}

void PQIVF::refresh_terms(const std::vector<int64_t> *which) {
    if (!have_codebooks || lists.empty()) return;
    int64_t max_len = 0;
    if (which) {
        if (which->empty()) return;
        for (int64_t l : *which) max_len = std::max(max_len, dev_len[l]);
        d_term_lists.resize(which->size());
        HANN_HIP(hipMemcpyAsync(d_term_lists.get(), which->data(), which->size() * sizeof(int64_t),
                                hipMemcpyHostToDevice, stream));
    } else {
        for (int64_t v : dev_len) max_len = std::max(max_len, v);
        // bounds start over with a full rebuild (entries only grow them otherwise)
        d_ymax_tmax.resize(2);
        HANN_HIP(hipMemsetAsync(d_ymax_tmax.get(), 0, 2 * sizeof(unsigned), stream));
    }
    launch_entry_terms(which ? d_term_lists.get() : nullptr, which ? (int64_t)which->size() : (int64_t)lists.size(),
                       max_len, d_list_start.get(), d_list_len.get(), d_codes.get(), code_bytes, d_has_codes.get(),
                       (int)dim, (int)m, (int)pmax, d_coarse.get(), d_cb.get(), d_term.get(), d_ymax_tmax.get(),
                       stream);
    // the host vector behind an asynchronous copy must outlive it
    HANN_HIP(hipStreamSynchronize(stream));
}

// --------------------------------------------------------------- search --
__global__ void k_min_i32(const int32_t *__restrict__ v, int64_t n, int32_t *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicMin(out, v[i]);
}

int PQIVF::device_min(const int32_t *v, int64_t n, hipStream_t st, Workspace &w) {
    int32_t *m = w.get<int32_t>(16, 1);
    HANN_HIP(hipMemsetAsync(m, 0x7F, sizeof(int32_t), st));   // 0x7F7F7F7F: above any nprobe
    hipLaunchKernelGGL(k_min_i32, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, v, n, m);
    HANN_LAUNCH_CHECK();
    int32_t r = 0;
    HANN_HIP(hipMemcpyAsync(&r, m, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HANN_HIP(hipStreamSynchronize(st));
    return r;
}

void PQIVF::search_device(const float *dq, int64_t nq, const int32_t *dk, const int32_t *dnp, int kmax, int pmax_probe,
                          const hann_search_params &p, float *out_key, int64_t *out_id, int32_t *out_cnt,
                          hipStream_t st, int min_nprobe, Workspace *wsp) {
    Workspace &w = wsp ? *wsp : ws;
    // the caller holds the index lock and has synced the device mirror
    if (nq == 0) return;
    const int d = (int)dim;
    const int K = (int)lists.size();
    int np = pmax_probe < K ? pmax_probe : K;
    if (np < 1) np = 1;
    const int R = p.rerank > 0 ? p.rerank : 0;
    // coarse step: the np nearest centroids by (rank, index); nearestCentroids
    // (pqivf/index.go:184-206). When every query probes every list no
    // selection is needed; otherwise the per-query prefix of the sorted
    // probe list is each query's set.
    std::array<hipEvent_t, 8> ev{};
    if (profile) {
        for (auto &e : ev) HANN_HIP(hipEventCreate(&e));
        HANN_HIP(hipEventRecord(ev[0], st));
        // [4] .. [7] start out recorded so that a search without a scan
        // kernel (or without a sample pass) measures zero
        for (int i = 4; i < 8; ++i) HANN_HIP(hipEventRecord(ev[i], st));
        kmark = &ev[4];
    }
    struct KmarkReset {
        hipEvent_t *&k;
        ~KmarkReset() { k = nullptr; }
    } kmark_reset{kmark};
    int64_t *probes;
    int probe_stride;
    double *probe_rank = nullptr;
    bool all_probe = np >= K;
    if (all_probe && dnp) {
        if (min_nprobe < 0) min_nprobe = device_min(dnp, nq, st, w);
        all_probe = min_nprobe >= K;
    }
    if (all_probe) {
        if ((int64_t)all_lists.n < (int64_t)K) {
            std::vector<int64_t> h((size_t)K);
            std::iota(h.begin(), h.end(), 0);
            all_lists.resize((size_t)K);
            HANN_HIP(hipMemcpyAsync(all_lists.get(), h.data(), h.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
            HANN_HIP(hipStreamSynchronize(st));
        }
        probes = all_lists.get();
        probe_stride = 0;
        np = K;
    } else {
        if (np > 512) fail(HANN_ERR_UNSUPPORTED, strf("nprobe %d above 512 (and below coarseK %d) is not supported", np, K));
        probes = w.get<int64_t>(5, (size_t)nq * np);
        probe_rank = w.get<double>(6, (size_t)nq * np);
        if (p.coarse_precision == HANN_PREC_BF16)
            // identical probes and ranks from the certified BF16 pass
            certified_topk(coarse_ctx, dq, nq, d_coarse.get(), kc, d, np, nullptr, 0, probes, probe_rank, st, dnp,
                           &coarse_bf16, &w.aux());
        else if (p.coarse_precision == HANN_PREC_FP32)
            launch_topk_exact(HANN_METRIC_SQEUCLIDEAN, dq, nq, d_coarse.get(), kc, d, np, probes, probe_rank, w, st);
        else
            fail(HANN_ERR_UNSUPPORTED, strf("coarse precision %d is not supported", p.coarse_precision));
        probe_stride = np;
    }
    const int selk = R > 0 ? kmax * R : kmax;
    const int cap = 2 * kp_for(selk);
    ScanArgs a{};
    a.queries = dq;
    a.d = d;
    a.m = (int)m;
    a.s = (int)(dim / m);
    a.pmax_cb = (int)(pmax > 0 ? pmax : 1);
    a.K = K;
    a.probes = probes;
    a.probe_stride = probe_stride;
    a.probe_cap = np;
    a.nprobe_q = dnp;
    a.nprobe_default = (int)nprobe;
    a.k_q = dk;
    a.k_default = kmax;
    a.rerank = R;
    a.allow_fallback = fallback ? 1 : 0;
    a.cap = cap;
    a.list_start = d_list_start.get();
    a.list_len = d_list_len.get();
    a.ids = d_ids.get();
    a.codes = d_codes.get();
    a.has_codes = d_has_codes.get();
    a.any_uncoded = any_uncoded ? 1 : 0;
    a.vecs = d_vecs.get();
    a.coarse = d_coarse.get();
    a.cb = d_cb.get();
    a.cb_len = d_cb_len.get();
    a.fb_counter = d_fb_counter.get();
    a.fell_back = nullptr;
    if (profile) HANN_HIP(hipEventRecord(ev[1], st));
    int mode = p.scan_mode;
    if (mode < SCAN_MODE_EXACT || mode > SCAN_MODE_LIST_FP8)
        fail(HANN_ERR_UNSUPPORTED, strf("scan mode %d is not supported", mode));
    bool list_scan = mode == SCAN_MODE_LIST_BF16 || mode == SCAN_MODE_LIST_BF16X3 || mode == SCAN_MODE_LIST_FP8;
    // the list-major scan's operands: 0 BF16, 1 split BF16, 3 FP8 (2 is the
    // exact scan)
    const int x3 = mode == SCAN_MODE_LIST_BF16X3 ? 1 : (mode == SCAN_MODE_LIST_FP8 ? 3 : 0);
    PQIVF::ListPlan lplan{};
    refresh_shape();
    if (list_scan) {
        // the list-major scan needs a rerank, codes for every entry, a probe
        // row per query, a supported shape and room for the batch's
        // buffers; otherwise the persistent table scan runs, and the first
        // time a reason holds for this index it is said on stderr
        const char *why = nullptr;
        if (R == 0)
            why = "it needs a rerank depth R > 0";
        else if (!have_codebooks || any_uncoded)
            why = "it needs PQ codes for every entry";
        else if (probe_stride == 0)
            why = "every query probes every list";
        else if ((int64_t)d_ids.n >= INT32_MAX || np > select_rows_max_probes())
            why = "the index or nprobe is too large for it";
        else if (reinterpret_cast<uintptr_t>(dq) % 16 != 0)
            why = "the queries are not 16-byte aligned";
        else if (!list_scan_supported((int)dim, (int)m, (int)pmax, code_bytes) ||
                 list_scan_lds_bytes_kind(x3, (int)m, (int)pmax, 2) > 160 * 1024 ||
                 select_records_lds_bytes() > 160 * 1024)
            why = "it covers 4-dimensional subvectors (d = 4 m, m in {8,16,24,32,48,64}) with uint8 codes";
        else if (!list_plan_fits(lplan = list_plan(nq, np, R > 0 ? kmax * R : kmax), nq))
            why = "the batch's buffers exceed its 32-bit addressing";
        if (why) {
            list_scan = false;
            mode = SCAN_MODE_ADC_F32;
            std::lock_guard<std::mutex> g(note_mu);
            if (list_notes.insert(why).second)
                std::fprintf(stderr, "hann_amd: the list-major scan was asked for but the table scan ran: %s (d %lld, m %lld)\n",
                             why, (long long)dim, (long long)m);
        }
    }
    if (list_scan) mode = SCAN_MODE_ADC_F32;   // the fallback scans
    if (!have_codebooks) {
        // no codebooks: every entry ranks by its exact vector
        a.any_uncoded = 1;
        mode = SCAN_MODE_ADC_F32;
    }
    // the plain path (the exact mode with a rerank) runs the list-major
    // exact scan where it covers the shape: the same distances bit for bit
    // and the same candidate rows; otherwise the per-query exact scan
    const bool exact_list = mode == SCAN_MODE_EXACT && R > 0 && have_codebooks && !any_uncoded && probe_stride > 0 &&
                            (int64_t)d_ids.n < INT32_MAX && np <= select_rows_max_probes() &&
                            exact_scan_supported((int)dim, (int)m, (int)pmax, code_bytes) &&
                            select_records_lds_bytes() <= 160 * 1024 &&
                            scan_lds_bytes(SCAN_MODE_EXACT, d, (int)m, (int)pmax, cap) <= 160 * 1024 &&
                            list_plan_fits(lplan = list_plan(nq, np, kmax * R, true), nq);
    const bool fused = (p.fused || list_scan) && have_codebooks && (mode == SCAN_MODE_ADC_F32 || mode == SCAN_MODE_ADC_F16);
    last_scan_mode = list_scan ? p.scan_mode : mode;
    last_scan_fused = exact_list ? 2 : ((fused && !list_scan) ? 1 : 0);
    auto scan = [&](const ScanArgs &args) {
        if (list_scan || exact_list) {
            const int xm = exact_list ? 2 : x3;
            const int nc = xm == 2 ? (int)(sizeof(kExactTunes) / sizeof(kExactTunes[0]))
                                   : (int)(sizeof(kListTunes) / sizeof(kListTunes[0]));
            if (p.list_config < 0 || p.list_config > nc)
                fail(HANN_ERR_INVALID_ARG, strf("list_config %d is outside 0 .. %d", p.list_config, nc));
            const ListTune t = p.list_config > 0 ? (xm == 2 ? kExactTunes : kListTunes)[p.list_config - 1]
                                                 : list_tuned(args, nq, xm, np, probe_rank, st, w, lplan);
            run_list_scan(args, nq, xm, np, probe_rank, st, w, lplan, t);
        }
        else if (fused)
            run_fused_scan(args, nq, np, probe_rank, st, w);
        else {
            mark_kernel(0, st);
            launch_scan(mode, code_bytes, args, nq, st);
            mark_kernel(1, st);
        }
    };
    if (R == 0) {
        a.out_key = out_key;
        a.out_id = out_id;
        a.out_cnt = out_cnt;
        a.out_stride = kmax;
        scan(a);
        if (profile) {
            HANN_HIP(hipEventRecord(ev[2], st));
            HANN_HIP(hipEventRecord(ev[3], st));
        }
    } else {
        float *ck = w.get<float>(7, (size_t)nq * selk);
        int64_t *cpos = w.get<int64_t>(8, (size_t)nq * selk);
        int32_t *ccnt = w.get<int32_t>(9, (size_t)nq);
        a.out_key = ck;
        a.out_id = cpos;
        a.out_cnt = ccnt;
        a.out_stride = selk;
        scan(a);
        if (profile) HANN_HIP(hipEventRecord(ev[2], st));
        launch_rerank(dq, d, nq, cpos, ccnt, selk, dk, kmax, kmax, d_vecs.get(), d_ids.get(), out_key, out_id, out_cnt,
                      kmax, st);
        if (profile) HANN_HIP(hipEventRecord(ev[3], st));
    }
    if (profile) t_events.push_back(ev);
}

void PQIVF::run_fused_scan(const ScanArgs &a, int64_t nq, int maxp, const double *probe_rank, hipStream_t st,
                           Workspace &w, bool main_scan) {
    if (num_cus == 0) {
        int dev = 0;
        HANN_HIP(hipGetDevice(&dev));
        HANN_HIP(hipDeviceGetAttribute(&num_cus, hipDeviceAttributeMultiprocessorCount, dev));
    }
    FusedArgs f{};
    f.term = d_term.get();
    f.probe_rank = a.probe_stride > 0 ? probe_rank : nullptr;
    f.order = w.get<int32_t>(12, (size_t)nq);
    f.counter = w.get<int32_t>(13, 1);
    // per-probe tables in LDS when each query has its own probe row
    f.lds_probes = a.probe_stride > 0 ? a.probe_cap : 0;
    // the batch's query tables in one pass (up to 512 MB of them at once)
    if ((size_t)nq * m * pmax * 4 <= (512ull << 20)) {
        float *S = w.get<float>(20, (size_t)nq * m * pmax);
        launch_query_tables(a.queries, nq, (int)dim, (int)m, (int)pmax, d_cb.get(), S, st);
        f.S = S;
    }
    const int kp = a.cap / 2;
    auto lds_for = [&](int wv) { return block_lds_bytes((int)dim, (int)m, (int)pmax, block_cap(kp, wv), f.lds_probes); };
    auto grid_for = [&](int wv) {
        const size_t shm = lds_for(wv);
        int per_cu = (int)std::max<size_t>(1, (160 * 1024) / std::max<size_t>(shm, 1));
        per_cu = std::min(per_cu, 32 / wv);
        return (int)std::min<int64_t>(nq, (int64_t)num_cus * per_cu);
    };
    auto fits = [&](int wv) { return lds_for(wv) <= 160 * 1024; };
    // the tuner: time every fitting waves-per-query choice once on this batch
    // shape, keep the fastest (results are identical for every choice)
    const uint64_t key = ((uint64_t)a.cap << 20) ^ (uint64_t)(maxp >= 64 ? 2 : maxp >= 16 ? 1 : 0);
    auto it = tuned_wpq.find(key);
    int wpq = 0;
    if (it != tuned_wpq.end()) {
        wpq = it->second;
    } else {
        float best = 1e30f;
        // trial runs count no brute-force fallbacks: the run below does
        ScanArgs ta = a;
        ta.fb_counter = nullptr;
        for (int wv : {2, 4, 8}) {
            if (!fits(wv)) continue;
            hipEvent_t e0, e1;
            HANN_HIP(hipEventCreate(&e0));
            HANN_HIP(hipEventCreate(&e1));
            HANN_HIP(hipEventRecord(e0, st));
            launch_scan_block(ta, f, nq, wv, kp, code_bytes, maxp, grid_for(wv), st);
            HANN_HIP(hipEventRecord(e1, st));
            HANN_HIP(hipEventSynchronize(e1));
            float ms = 0;
            HANN_HIP(hipEventElapsedTime(&ms, e0, e1));
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            if (ms < best) {
                best = ms;
                wpq = wv;
            }
        }
        if (wpq == 0) fail(HANN_ERR_UNSUPPORTED, "persistent scan: no configuration fits in LDS");
        tuned_wpq[key] = wpq;
        last_tune = strf("persistent scan tuned: waves/query %d (%.3f ms)", wpq, best);
    }
    if (main_scan) mark_kernel(0, st);
    launch_scan_block(a, f, nq, wpq, kp, code_bytes, maxp, grid_for(wpq), st);
    if (main_scan) mark_kernel(1, st);
}

// The list-major scan's batch buffers: the arena's layout (one workspace
// slot, the buffers that start at zero first so one fill clears them), the
// record room per query and the sample rows' size.
struct ListArena {
    size_t off = 0;
    size_t o_lcnt, o_lscnt, o_iord, o_ctr, o_parts, o_parts_f, o_qcnt, o_zero_end, o_lstart, o_ioff, o_lfill, o_lbfill,
        o_qsnp, o_qsamp, o_tot, o_qtot, o_soff, o_theta, o_qscale, o_ovf, o_pq, o_pqc, o_psc, o_poff, o_items,
        o_items_s, o_pkey;
    size_t carve(size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    }
    ListArena(int64_t K, int64_t nq, int64_t pairs_max, int64_t items_max, int selk) {
        o_lcnt = carve((size_t)K * 4);
        o_lscnt = carve((size_t)K * 4);
        o_iord = carve((size_t)list_item_order_ints() * 4);
        o_ctr = carve(16);
        o_parts = carve(kListScanParts * 4);
        o_parts_f = carve(kListScanParts * 4);
        o_qcnt = carve((size_t)nq * 4);
        o_zero_end = off;
        o_lstart = carve((size_t)K * 4);
        o_ioff = carve((size_t)K * 4);
        o_lfill = carve((size_t)K * 4);
        o_lbfill = carve((size_t)K * 4);
        o_qsnp = carve((size_t)nq * 4);
        o_qsamp = carve((size_t)nq * 8);
        o_tot = carve(16);
        o_qtot = carve((size_t)nq * 8);
        o_soff = carve((size_t)nq * 8);
        o_theta = carve((size_t)nq * 8);
        o_qscale = carve((size_t)nq * 4);
        o_ovf = carve((size_t)nq * 4);
        o_pq = carve((size_t)pairs_max * 4);
        o_pqc = carve((size_t)pairs_max * 4);
        o_psc = carve((size_t)pairs_max * 4);
        o_poff = carve((size_t)pairs_max * 8);
        o_items = carve((size_t)items_max * 16);
        o_items_s = carve((size_t)items_max * 16);
        o_pkey = carve((size_t)nq * selk * 4);
    }
};

void PQIVF::refresh_shape() const {
    if (shape_epoch.load(std::memory_order_acquire) == mirror_epoch) return;
    int64_t max_len = 0, chunks[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t v : dev_len) {
        max_len = std::max(max_len, v);
        for (int c = 0; c < 6; ++c) chunks[c] += ceil_div(v, int64_t(1) << (6 + c));
    }
    shape_max_len.store(max_len, std::memory_order_relaxed);
    for (int c = 0; c < 6; ++c) shape_chunks[c].store(chunks[c], std::memory_order_relaxed);
    shape_epoch.store(mirror_epoch, std::memory_order_release);
}

PQIVF::ListPlan PQIVF::list_plan(int64_t nq, int np, int selk, bool exact) const {
    ListPlan L;
    // work items: a list with p pairs and n entries has ceil(p / G) *
    // ceil(n / C) of them (G queries, C entries an item), at most
    // (p / G + 1) * ceil(n / C)
    // sized for the smallest chunk and query group the tuner may pick
    const ItemShape sh = exact ? kExactItemsMin : kListItemsMin;
    const int64_t chunk = int64_t(1) << sh.ch, group = int64_t(1) << sh.qg;
    const int64_t max_len = shape_max_len.load(std::memory_order_relaxed);
    const int64_t chunks_all = shape_chunks[sh.ch - 6].load(std::memory_order_relaxed);
    L.exact = exact;
    L.pairs_max = nq * (int64_t)np;
    L.items_max = ceil_div(L.pairs_max, group) * ceil_div(max_len, chunk) + chunks_all + 1;
    // a query's sample: at least one probe, then on while fewer than
    // 4 selk entries come before: at most 4 selk + max_len entries (the
    // bound below also covers a two-probe minimum), each probe's run
    // rounded up to 8 slots
    L.sample_slots = nq * (std::max<int64_t>(4 * (int64_t)selk + max_len, 2 * max_len) + 7 * (int64_t)np);
    // record room: the filter pass keeps about 1.5 selk records a query
    // (at most every probed entry); a query with more takes the table scan
    L.capq = std::max<int64_t>(8 * (int64_t)selk, 256);
    L.arena = ListArena((int64_t)lists.size(), nq, L.pairs_max, L.items_max, selk).off;
    return L;
}

ListTune PQIVF::list_tuned(const ScanArgs &a, int64_t nq, int x3, int maxp, const double *probe_rank, hipStream_t st,
                           Workspace &w, const ListPlan &L) {
    const uint64_t key = ((uint64_t)x3 << 40) ^ ((uint64_t)a.out_stride << 8) ^
                         (uint64_t)(maxp >= 64 ? 2 : maxp >= 16 ? 1 : 0);
    auto it = tuned_list.find(key);
    if (it != tuned_list.end()) return it->second;
    const ListTune *cands = x3 == 2 ? kExactTunes : kListTunes;
    const int nc = x3 == 2 ? (int)(sizeof(kExactTunes) / sizeof(kExactTunes[0]))
                           : (int)(sizeof(kListTunes) / sizeof(kListTunes[0]));
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HANN_HIP(hipStreamIsCapturing(st, &cs));
    if (cs != hipStreamCaptureStatusNone) return cands[0];
    // trial runs count no brute-force fallbacks: the caller's run does
    ScanArgs ta = a;
    ta.fb_counter = nullptr;
    // one untimed run first (the kernels' first launches load code)
    run_list_scan(ta, nq, x3, maxp, probe_rank, st, w, L, cands[0]);
    float best = 1e30f;
    int bi = 0;
    for (int i = 0; i < nc; ++i) {
        hipEvent_t e0, e1;
        HANN_HIP(hipEventCreate(&e0));
        HANN_HIP(hipEventCreate(&e1));
        HANN_HIP(hipEventRecord(e0, st));
        run_list_scan(ta, nq, x3, maxp, probe_rank, st, w, L, cands[i]);
        HANN_HIP(hipEventRecord(e1, st));
        HANN_HIP(hipEventSynchronize(e1));
        float ms = 0;
        HANN_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (ms < best) {
            best = ms;
            bi = i;
        }
    }
    tuned_list[key] = cands[bi];
    last_tune = strf("list-major %s scan tuned: %d queries x %d entries per item, %d partitions (%.3f ms)",
                     x3 == 2 ? "exact" : "MFMA", 1 << cands[bi].sh.qg, 1 << cands[bi].sh.ch, cands[bi].parts, best);
    return cands[bi];
}

void PQIVF::run_list_scan(const ScanArgs &a, int64_t nq, int x3, int maxp, const double *probe_rank,
                          hipStream_t st, Workspace &w, const ListPlan &L, const ListTune &t) {
    const int K = (int)lists.size();
    const int selk = a.out_stride;
    const int kp = a.cap / 2;
    const ListArena A(K, nq, L.pairs_max, L.items_max, selk);
    char *base = w.get<char>(17, A.off);
    auto at = [&](size_t o) { return base + o; };
    int32_t *lcnt = reinterpret_cast<int32_t *>(at(A.o_lcnt));
    int32_t *lstart = reinterpret_cast<int32_t *>(at(A.o_lstart));
    int32_t *ioff = reinterpret_cast<int32_t *>(at(A.o_ioff));
    int32_t *lfill = reinterpret_cast<int32_t *>(at(A.o_lfill));
    int32_t *lscnt = reinterpret_cast<int32_t *>(at(A.o_lscnt));   // sample pairs per list
    int32_t *lbfill = reinterpret_cast<int32_t *>(at(A.o_lbfill));
    int32_t *qsamp_np = reinterpret_cast<int32_t *>(at(A.o_qsnp));   // probes in a query's sample
    int64_t *qsamp = reinterpret_cast<int64_t *>(at(A.o_qsamp));     // row slots of a query's sample
    int32_t *totals = reinterpret_cast<int32_t *>(at(A.o_tot));
    int64_t *qtot = reinterpret_cast<int64_t *>(at(A.o_qtot));
    int64_t *soff = reinterpret_cast<int64_t *>(at(A.o_soff));   // sample row starts
    uint32_t *qcnt = reinterpret_cast<uint32_t *>(at(A.o_qcnt));   // records per query
    int32_t *ovf = reinterpret_cast<int32_t *>(at(A.o_ovf));
    int32_t *ctr = reinterpret_cast<int32_t *>(at(A.o_ctr));   // [1] fallbacks, [2] fallback queue
    int32_t *parts = reinterpret_cast<int32_t *>(at(A.o_parts));   // items taken per partition (sample pass)
    int32_t *parts_f = reinterpret_cast<int32_t *>(at(A.o_parts_f));   // (filter pass)
    int32_t *pair_q = reinterpret_cast<int32_t *>(at(A.o_pq));
    float *pair_qc = reinterpret_cast<float *>(at(A.o_pqc));
    float *pair_scale = reinterpret_cast<float *>(at(A.o_psc));
    int64_t *pair_off = reinterpret_cast<int64_t *>(at(A.o_poff));
    int4 *items = reinterpret_cast<int4 *>(at(A.o_items));
    int4 *items_s = reinterpret_cast<int4 *>(at(A.o_items_s));   // the sample pass's items
    float *pkey = reinterpret_cast<float *>(at(A.o_pkey));
    const bool exact = x3 == 2;
    float *q_scale = exact ? reinterpret_cast<float *>(at(A.o_qscale)) : nullptr;   // key scale per query
    const ItemShape sh = t.sh;
    if (num_cus == 0) {
        int dev = 0;
        HANN_HIP(hipGetDevice(&dev));
        HANN_HIP(hipDeviceGetAttribute(&num_cus, hipDeviceAttributeMultiprocessorCount, dev));
    }
    const size_t shm = exact ? exact_scan_lds_bytes((int)dim, (int)m, (int)pmax)
                             : list_scan_lds_bytes_kind(x3, (int)m, (int)pmax, 2);
    const int grid = num_cus * (int)std::max<size_t>(1, std::min<size_t>(4, (160 * 1024) / shm));
    // the filter pass's records, capq a query; the sample pass's keys
    uint64_t *rec = w.get<uint64_t>(18, (size_t)(nq * L.capq));
    uint16_t *srow = w.get<uint16_t>(19, (size_t)L.sample_slots);
    uint32_t *theta = reinterpret_cast<uint32_t *>(at(A.o_theta));   // [nq]

    // 1. the batch's (query, list) pairs grouped by list, as work items (and
    //    the sample pass's over the sample pairs), and each query's sample row
    launch_fill_i32(lcnt, (int64_t)(A.o_zero_end - A.o_lcnt) / 4, 0, st);
    launch_pair_count(a, nq, lcnt, lscnt, qtot, qsamp_np, qsamp, st);
    launch_row_offsets(qsamp, nq, soff, st);
    launch_pair_offsets(lcnt, d_list_len.get(), K, lstart, ioff, lfill, lbfill, totals, sh, st);
    launch_pair_scatter(a, nq, probe_rank, d_ymax_tmax.get(), lstart, lcnt, qsamp_np, soff, lfill, lbfill, pair_q,
                        pair_qc, pair_scale, pair_off, q_scale, st);
    launch_make_items(lcnt, lscnt, lstart, d_list_len.get(), K, reinterpret_cast<int32_t *>(at(A.o_iord)), items,
                      items_s, totals + 2, sh, st);
    ListScanArgs la{};
    la.queries = a.queries;
    la.d = (int)dim;
    la.m = (int)m;
    la.s = (int)(dim / m);
    la.P = (int)pmax;
    la.dk = (int)ceil_div(dim, 16) * 16;
    la.nq = nq;
    la.list_start = d_list_start.get();
    la.list_len = d_list_len.get();
    la.codes = d_codes.get();
    la.term = d_term.get();
    la.cb = d_cb.get();
    la.pair_q = pair_q;
    la.pair_qc = pair_qc;
    la.pair_scale = pair_scale;
    la.pair_off = pair_off;
    la.items = items;
    la.n_items = totals + 1;
    la.item_counter = parts;
    la.n_parts = std::min(std::min(t.parts, kListScanParts), grid);
    la.dist = srow;
    la.dist_bytes = (uint32_t)(L.sample_slots * 2);
    la.theta = theta;
    la.rec = rec;
    la.qcnt = qcnt;
    la.capq = (uint32_t)L.capq;
    la.coarse = d_coarse.get();
    la.chunk = 1 << sh.ch;
    la.cb_scale = cb_fp8_scale;
    // 2. the sample pass: every key of each query's nearest probed lists;
    //    then each query's bound theta, the (k R)-th smallest of them
    {
        ListScanArgs ls = la;
        ls.items = items_s;
        ls.n_items = totals + 2;
        mark_kernel(2, st);
        if (exact)
            launch_scan_exact_kind(ls, 1, grid, st);
        else
            launch_scan_list_kind(ls, x3, 1, grid, st);
        mark_kernel(3, st);
    }
    launch_theta(a, nq, srow, soff, qtot, qsamp, theta, st);
    // 3. the filter pass: every (query, entry) key, and the records of those
    //    <= theta in the query's record room
    la.item_counter = parts_f;
    mark_kernel(0, st);
    if (exact)
        launch_scan_exact_kind(la, 2, grid, st);
    else
        launch_scan_list_kind(la, x3, 2, grid, st);
    mark_kernel(1, st);
    // 4. the selk best records per query, as the rerank's rows
    launch_select_records(a, nq, rec, qcnt, (uint32_t)L.capq, qtot, a.out_id, a.out_cnt, selk, ovf, ctr + 1, q_scale,
                          st);
    // 5. the queries the select passed on: the table scan ranks them (over
    //    all lists where the reference's brute-force rule applies, counter
    //    included; over their probes when their records overflowed)
    if (exact) {
        // the exact scan's passed-on queries: the per-query exact scan
        ScanArgs fa = a;
        fa.out_key = pkey;
        launch_scan_ordered(SCAN_MODE_EXACT, code_bytes, fa, nq, ovf, ctr + 1, st);
        return;
    }
    FusedArgs f{};
    f.term = d_term.get();
    f.probe_rank = probe_rank;
    f.order = ovf;
    f.counter = ctr + 2;
    f.nq_dev = ctr + 1;
    f.lds_probes = a.probe_cap;
    ScanArgs fa = a;
    fa.out_key = pkey;
    const int wv = 4;
    const size_t fshm = block_lds_bytes((int)dim, (int)m, (int)pmax, block_cap(kp, wv), f.lds_probes);
    if (fshm > 160 * 1024) fail(HANN_ERR_UNSUPPORTED, "list-major scan: the fallback scan does not fit in LDS");
    // fallbacks are rare: one workgroup per CU drains the queue
    const int fgrid = (int)std::min<int64_t>(nq, (int64_t)num_cus);
    launch_scan_block(fa, f, nq, wv, kp, code_bytes, maxp, fgrid, st);
}

int64_t PQIVF::search_scratch_bytes(int64_t nq, int kmax, int pmax_probe, const hann_search_params &p) const {
    if (nq <= 0) return 0;
    const int64_t K = (int64_t)lists.size();
    const int np = (int)std::max<int64_t>(1, std::min<int64_t>(pmax_probe, K));
    const int R = p.rerank > 0 ? p.rerank : 0;
    const int selk = R > 0 ? kmax * R : kmax;
    int64_t b = 0;
    auto add = [&](int64_t bytes) { b += bytes + 256; };   // the workspace's slack a slot
    add(4);   // device_min
    if (np < K) {
        add(nq * np * 8);   // probes
        add(nq * np * 8);   // probe ranks
        if (p.coarse_precision == HANN_PREC_BF16)
            b += (int64_t)certified_topk_scratch_bytes(nq, kc, (int)dim, np, true);
        else
            b += (int64_t)topk_exact_scratch_bytes(HANN_METRIC_SQEUCLIDEAN, nq, kc, np);
    }
    if (R > 0) {
        add(nq * selk * 4);   // candidate keys
        add(nq * selk * 8);   // candidate positions
        add(nq * 4);
    }
    // the table scan (also the list-major scan's fallback)
    add(nq * 4);
    add(4);
    if (nq * m * pmax * 4 <= (int64_t(512) << 20)) add(nq * m * pmax * 4);
    const bool lm = p.scan_mode == SCAN_MODE_LIST_BF16 || p.scan_mode == SCAN_MODE_LIST_BF16X3 ||
                    p.scan_mode == SCAN_MODE_LIST_FP8;
    if (lm || (p.scan_mode == SCAN_MODE_EXACT && R > 0)) {
        const ListPlan L = list_plan(nq, np, selk, !lm);
        add((int64_t)L.arena);
        add(nq * L.capq * 8);
        add(L.sample_slots * 2);
    }
    return b;
}

bool PQIVF::list_plan_fits(const ListPlan &L, int64_t nq) {
    // 16-bit sample keys addressed below 2^32 - 64 bytes; record rooms and
    // entry positions addressed in 32 bits
    return L.sample_slots * 2 < (int64_t(1) << 32) - 64 && nq * L.capq < (int64_t(1) << 31) &&
           L.pairs_max < (int64_t(1) << 31);
}

void PQIVF::resolve_timers() {
    for (auto &ev : t_events) {
        HANN_HIP(hipEventSynchronize(ev[3]));
        float a = 0, b = 0, c = 0;
        HANN_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
        HANN_HIP(hipEventElapsedTime(&b, ev[1], ev[2]));
        HANN_HIP(hipEventElapsedTime(&c, ev[2], ev[3]));
        float kt = 0, ks = 0;
        HANN_HIP(hipEventElapsedTime(&kt, ev[4], ev[5]));
        HANN_HIP(hipEventElapsedTime(&ks, ev[6], ev[7]));
        t_coarse_ms += a;
        t_scan_ms += b;
        t_rerank_ms += c;
        t_kernel_ms += kt;
        t_sample_ms += ks;
        ++t_calls;
        for (auto &e : ev) (void)hipEventDestroy(e);
    }
    t_events.clear();
}

void PQIVF::check_searchable(int64_t k, size_t len) const {
    // Search's validation order (pqivf/index.go:854-872)
    if (k <= 0) fail(HANN_ERR_INVALID_ARG, "k must be positive");
    if (!trained)
        fail(HANN_ERR_NOT_TRAINED, "PQIVF index is not trained. Call Train() after adding data before searching");
    if ((int64_t)len != dim)
        fail(HANN_ERR_INVALID_ARG,
             strf("query dimension %zu does not match index dimension %lld", len, (long long)dim));
    bool any = false;
    for (const ListData &L : lists) any = any || L.exists;
    if (!any) fail(HANN_ERR_EMPTY, "index is empty");
}

void PQIVF::search_host(const float *queries, int64_t nq, const int32_t *k, const int32_t *np,
                        const hann_search_params &p, int64_t *out_ids, double *out_dist, int32_t *out_counts) {
    int kmax = 0, pmx = (int)nprobe;
    for (int64_t q = 0; q < nq; ++q) {
        kmax = std::max(kmax, k ? k[q] : 0);
        if (np) pmx = std::max(pmx, np[q]);
    }
    if (np) {
        pmx = 0;
        for (int64_t q = 0; q < nq; ++q) pmx = std::max(pmx, np[q]);
    }
    for (int64_t q = 0; q < nq; ++q) check_searchable(k ? k[q] : 0, (size_t)dim);
    if (nq == 0) return;
    sync_device();
    hipStream_t st = stream;
    float *dq = ws.get<float>(0, (size_t)nq * dim);
    int32_t *dk = ws.get<int32_t>(1, (size_t)nq);
    int32_t *dn = np ? ws.get<int32_t>(2, (size_t)nq) : nullptr;
    float *ok = ws.get<float>(3, (size_t)nq * kmax);
    int64_t *oi = ws.get<int64_t>(4, (size_t)nq * kmax);
    int32_t *oc = ws.get<int32_t>(10, (size_t)nq);
    HANN_HIP(hipMemcpyAsync(dq, queries, (size_t)nq * dim * sizeof(float), hipMemcpyHostToDevice, st));
    HANN_HIP(hipMemcpyAsync(dk, k, (size_t)nq * sizeof(int32_t), hipMemcpyHostToDevice, st));
    if (np) HANN_HIP(hipMemcpyAsync(dn, np, (size_t)nq * sizeof(int32_t), hipMemcpyHostToDevice, st));
    int pmin = np ? INT32_MAX : (int)nprobe;
    if (np)
        for (int64_t q = 0; q < nq; ++q) pmin = std::min(pmin, np[q]);
    search_device(dq, nq, dk, dn, kmax, pmx, p, ok, oi, oc, st, pmin);
    std::vector<float> hk((size_t)nq * kmax);
    HANN_HIP(hipMemcpyAsync(hk.data(), ok, hk.size() * sizeof(float), hipMemcpyDeviceToHost, st));
    HANN_HIP(hipMemcpyAsync(out_ids, oi, (size_t)nq * kmax * sizeof(int64_t), hipMemcpyDeviceToHost, st));
    HANN_HIP(hipMemcpyAsync(out_counts, oc, (size_t)nq * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HANN_HIP(hipStreamSynchronize(st));
    // FromRank: float64 square root of the widened float32 rank
    // (pqivf/index.go:945-947, core/metric.go:156)
    for (size_t i = 0; i < hk.size(); ++i) out_dist[i] = std::sqrt((double)hk[i]);
}

int64_t PQIVF::fallback_count() {
    int64_t devc = 0;
    if (stream && d_fb_counter.get()) {
        unsigned long long v = 0;
        HANN_HIP(hipMemcpyAsync(&v, d_fb_counter.get(), sizeof(v), hipMemcpyDeviceToHost, stream));
        HANN_HIP(hipStreamSynchronize(stream));
        devc = (int64_t)v;
    }
    return devc;
}

int64_t PQIVF::count() const {
    int64_t c = (int64_t)pending.size();
    for (const ListData &L : lists) c += (int64_t)L.ids.size();
    return c;
}

// ------------------------------------------------------------ save/load --
namespace {
constexpr char kMagic[8] = {'H', 'A', 'N', 'N', 'P', 'Q', 'I', 'V'};
constexpr uint32_t kFormatVersion = 1;

struct Writer {
    std::vector<unsigned char> b;
    template <typename T>
    void put(const T &v) {
        const unsigned char *p = reinterpret_cast<const unsigned char *>(&v);
        b.insert(b.end(), p, p + sizeof(T));
    }
    void raw(const void *p, size_t n) {
        const unsigned char *c = reinterpret_cast<const unsigned char *>(p);
        b.insert(b.end(), c, c + n);
    }
};

struct Reader {
    const unsigned char *p;
    size_t n, at = 0;
    template <typename T>
    T get() {
        if (at + sizeof(T) > n) fail(HANN_ERR_FORMAT, "serialized index is truncated");
        T v;
        std::memcpy(&v, p + at, sizeof(T));
        at += sizeof(T);
        return v;
    }
    void raw(void *dst, size_t bytes) {
        if (at + bytes > n || bytes > n) fail(HANN_ERR_FORMAT, "serialized index is truncated");
        std::memcpy(dst, p + at, bytes);
        at += bytes;
    }
    int64_t count(size_t elem) {
        int64_t c = get<int64_t>();
        if (c < 0 || (elem && (uint64_t)c > (n - at) / elem + 1)) fail(HANN_ERR_FORMAT, "serialized index is corrupt");
        return c;
    }
};
}  // namespace

std::vector<unsigned char> PQIVF::save() const {
    // serializedPQIVF (pqivf/index.go:969-984) in a binary layout
    Writer w;
    w.raw(kMagic, 8);
    w.put<uint32_t>(kFormatVersion);
    w.put<int64_t>(dim);
    w.put<int64_t>(coarse_k);
    w.put<int64_t>(m);
    w.put<int64_t>(pq_k);
    w.put<int64_t>(iters);
    w.put<int64_t>(nprobe);
    w.put<uint8_t>(fallback ? 1 : 0);
    w.put<uint8_t>(trained ? 1 : 0);
    w.put<uint64_t>(seed);
    w.put<uint64_t>(train_epoch);
    w.put<int64_t>(kc);
    w.raw(coarse.data(), coarse.size() * sizeof(float));
    w.put<uint8_t>(have_codebooks ? 1 : 0);
    w.put<int64_t>(pmax);
    for (int64_t i = 0; i < m; ++i) w.put<int64_t>(i < (int64_t)cb_len.size() ? cb_len[i] : 0);
    w.raw(codebooks.data(), codebooks.size() * sizeof(float));
    int64_t nl = 0;
    for (const ListData &L : lists) nl += L.exists ? 1 : 0;
    w.put<int64_t>(nl);
    for (size_t c = 0; c < lists.size(); ++c) {
        const ListData &L = lists[c];
        if (!L.exists) continue;
        w.put<int64_t>((int64_t)c);
        w.put<int64_t>((int64_t)L.ids.size());
        w.raw(L.ids.data(), L.ids.size() * sizeof(int64_t));
        w.raw(L.has_codes.data(), L.has_codes.size());
        w.raw(L.codes.data(), L.codes.size() * sizeof(int32_t));
        w.raw(L.vecs.data(), L.vecs.size() * sizeof(float));
    }
    w.put<int64_t>((int64_t)pending.size());
    for (size_t i = 0; i < pending.size(); ++i) {
        w.put<int64_t>(pending.ids[i]);
        w.put<int64_t>(dim);
        w.raw(&pending.vecs[i * dim], dim * sizeof(float));
    }
    return std::move(w.b);
}

void PQIVF::load(const void *buf, size_t len) {
    // GobDecode (pqivf/index.go:1020-1084): version and parameter checks,
    // then the new state replaces the old one
    Reader r{reinterpret_cast<const unsigned char *>(buf), len};
    char magic[8];
    r.raw(magic, 8);
    if (std::memcmp(magic, kMagic, 8) != 0) fail(HANN_ERR_FORMAT, "not a hann PQIVF index file");
    const uint32_t ver = r.get<uint32_t>();
    if (ver > kFormatVersion)
        fail(HANN_ERR_FORMAT, strf("index file has format version %u, but this build supports up to %u", ver, kFormatVersion));
    const int64_t ddim = r.get<int64_t>(), dck = r.get<int64_t>(), dm = r.get<int64_t>(), dpk = r.get<int64_t>();
    const int64_t dit = r.get<int64_t>(), dnp = r.get<int64_t>();
    const bool dfb = r.get<uint8_t>() != 0, dtr = r.get<uint8_t>() != 0;
    const uint64_t dseed = r.get<uint64_t>(), depoch = r.get<uint64_t>();
    if (ddim <= 0) fail(HANN_ERR_FORMAT, strf("serialized index has invalid dimension %lld", (long long)ddim));
    if (dm <= 0 || ddim % dm != 0)
        fail(HANN_ERR_FORMAT, strf("serialized index has invalid subquantizer count %lld for dimension %lld", (long long)dm,
                                   (long long)ddim));
    if (dck <= 0 || dpk <= 0 || dit <= 0)
        fail(HANN_ERR_FORMAT, strf("serialized index has invalid training parameters: coarseK %lld, pqK %lld, kMeansIters %lld",
                                   (long long)dck, (long long)dpk, (long long)dit));
    const int64_t dkc = r.count((size_t)ddim * 4);
    std::vector<float> C((size_t)(dkc * ddim));
    r.raw(C.data(), C.size() * sizeof(float));
    const bool dhave = r.get<uint8_t>() != 0;
    const int64_t dpmax = r.count(0);
    std::vector<int64_t> cbl((size_t)dm);
    r.raw(cbl.data(), cbl.size() * sizeof(int64_t));
    const int64_t s = ddim / dm;
    if (dpmax > (int64_t)((len - r.at) / 4 / std::max<int64_t>(1, dm * s) + 1)) fail(HANN_ERR_FORMAT, "serialized index is corrupt");
    std::vector<float> CB((size_t)(dm * dpmax * s));
    r.raw(CB.data(), CB.size() * sizeof(float));
    for (int64_t i = 0; i < dm; ++i)
        if (cbl[i] < 0 || cbl[i] > dpmax) fail(HANN_ERR_FORMAT, "serialized index has an invalid codebook length");
    std::vector<ListData> L((size_t)dkc);
    std::unordered_map<int64_t, int64_t> i2c;
    const int64_t nl = r.count(16);
    for (int64_t j = 0; j < nl; ++j) {
        const int64_t c = r.get<int64_t>();
        const int64_t n = r.count((size_t)(8 + 1 + 4 * dm + 4 * ddim));
        if (c < 0 || c >= dkc) fail(HANN_ERR_FORMAT, strf("serialized list %lld has no coarse centroid", (long long)c));
        ListData &x = L[c];
        x.exists = true;
        x.ids.reserve(n + std::max<int64_t>(8, n / 8));
        x.has_codes.reserve(n + std::max<int64_t>(8, n / 8));
        x.codes.reserve((n + std::max<int64_t>(8, n / 8)) * dm);
        x.vecs.reserve((n + std::max<int64_t>(8, n / 8)) * ddim);
        x.ids.resize(n);
        x.has_codes.resize(n);
        x.codes.resize(n * dm);
        x.vecs.resize(n * ddim);
        r.raw(x.ids.data(), n * 8);
        r.raw(x.has_codes.data(), n);
        r.raw(x.codes.data(), n * dm * 4);
        r.raw(x.vecs.data(), n * ddim * 4);
        for (int64_t id : x.ids) i2c[id] = c;
    }
    PendingStore pend;
    pend.dim = ddim;
    const int64_t npend = r.count(16);
    std::vector<float> v((size_t)ddim);
    for (int64_t j = 0; j < npend; ++j) {
        const int64_t id = r.get<int64_t>();
        const int64_t n = r.count(4);
        if (n != ddim) fail(HANN_ERR_FORMAT, strf("pending vector %lld has %lld values, expected %lld", (long long)id,
                                                  (long long)n, (long long)ddim));
        r.raw(v.data(), n * 4);
        if (!pend.count(id)) pend.put(id, v.data());
    }
    IndexState st;
    st.dim = ddim;
    st.coarse_k = dck;
    st.m = dm;
    st.pq_k = dpk;
    st.iters = dit;
    st.nprobe = dnp;
    st.fallback = dfb;
    st.trained = dtr;
    st.has_seed = true;
    st.seed = dseed;
    st.train_epoch = depoch;
    st.kc = dkc;
    st.coarse.swap(C);
    st.have_codebooks = dhave;
    st.pmax = dpmax;
    st.cb_len.swap(cbl);
    st.codebooks.swap(CB);
    st.lists.swap(L);
    st.id_to_cluster.swap(i2c);
    std::swap(st.pending, pend);
    install(std::move(st));
}

void PQIVF::install(IndexState &&st) {
    dim = st.dim;
    coarse_k = st.coarse_k;
    m = st.m;
    pq_k = st.pq_k;
    iters = st.iters;
    // a file without a probe count keeps the configured one, else 3
    // (pqivf/index.go:1070-1074)
    if (st.nprobe > 0) nprobe = st.nprobe;
    else if (nprobe <= 0) nprobe = 3;
    fallback = st.fallback;
    trained = st.trained;
    if (st.has_seed) {
        seed = st.seed;
        train_epoch = st.train_epoch;
    }
    kc = st.kc;
    coarse.swap(st.coarse);
    have_codebooks = st.have_codebooks;
    pmax = st.pmax;
    cb_len.swap(st.cb_len);
    codebooks.swap(st.codebooks);
    lists.swap(st.lists);
    id_to_cluster.swap(st.id_to_cluster);
    std::swap(pending, st.pending);
    mark_dirty();
}

}  // namespace hann

// =================================================================== C-ABI
using namespace hann;

extern "C" {

void hann_pqivf_default_options(hann_pqivf_options *o) {
    std::memset(o, 0, sizeof(*o));
    o->bruteforce_fallback = 1;
    o->seed = 0;
}

int hann_pqivf_new(int64_t dim, const hann_pqivf_options *opts, hann_pqivf **out) {
    HANN_API_BEGIN
    *out = nullptr;
    // pqivf.New (pqivf/index.go:106-146)
    if (dim <= 0) fail(HANN_ERR_INVALID_ARG, strf("dimension (%lld) must be positive", (long long)dim));
    hann_pqivf_options o;
    hann_pqivf_default_options(&o);
    if (opts) o = *opts;
    auto pick = [&](int64_t v, int bit, int64_t def) { return (v != 0 || (o.explicit_mask >> bit) & 1) ? v : def; };
    auto h = std::make_unique<hann_pqivf>();
    PQIVF &x = h->ix;
    x.dim = dim;
    x.pending.dim = dim;
    x.coarse_k = pick(o.coarse_k, 0, 16);
    x.m = pick(o.num_subquantizers, 1, 0);
    if (x.m == 0 && !((o.explicit_mask >> 1) & 1)) x.m = default_num_subquantizers(dim);
    x.pq_k = pick(o.pq_k, 2, 16);
    x.iters = pick(o.kmeans_iters, 3, 10);
    x.nprobe = pick(o.candidate_clusters, 4, 3);
    x.fallback = ((o.explicit_mask >> 5) & 1) ? o.bruteforce_fallback != 0 : true;
    x.seed = o.seed;
    x.train_precision = o.train_precision;
    if (o.kmeans_init != 0 && o.kmeans_init != 1)
        fail(HANN_ERR_UNSUPPORTED, strf("k-means init %d is not supported", (int)o.kmeans_init));
    x.kmeans_init = o.kmeans_init;
    x.coarse_sample = o.train_sample;
    x.pq_sample = o.pq_train_sample;
    if (x.coarse_k <= 0) fail(HANN_ERR_INVALID_ARG, strf("coarseK (%lld) must be positive", (long long)x.coarse_k));
    if (x.pq_k <= 0) fail(HANN_ERR_INVALID_ARG, strf("pqK (%lld) must be positive", (long long)x.pq_k));
    if (x.iters <= 0) fail(HANN_ERR_INVALID_ARG, strf("kMeansIters (%lld) must be positive", (long long)x.iters));
    if (x.m <= 0) fail(HANN_ERR_INVALID_ARG, strf("numSubquantizers (%lld) must be positive", (long long)x.m));
    if (dim % x.m != 0)
        fail(HANN_ERR_INVALID_ARG, strf("dimension (%lld) must be divisible by numSubquantizers (%lld)", (long long)dim,
                                        (long long)x.m));
    *out = h.release();
    HANN_API_END
}

void hann_pqivf_free(hann_pqivf *idx) { delete idx; }

#define IDX_CHECK()                                                    \
    if (!idx) fail(HANN_ERR_INVALID_ARG, "index handle is null");      \
    PQIVF &x = idx->ix

int hann_pqivf_add(hann_pqivf *idx, int64_t id, const float *vec, size_t len) {
    HANN_API_BEGIN
    IDX_CHECK();
    std::unique_lock<std::shared_mutex> lk(x.mu);
    x.add_locked(id, vec, len);
    HANN_API_END
}

int hann_pqivf_bulk_add(hann_pqivf *idx, const int64_t *ids, const float *flat, const size_t *lens, size_t n) {
    HANN_API_BEGIN
    IDX_CHECK();
    std::unique_lock<std::shared_mutex> lk(x.mu);
    x.bulk_add(ids, flat, lens, n);
    HANN_API_END
}

int hann_pqivf_delete(hann_pqivf *idx, int64_t id) {
    HANN_API_BEGIN
    IDX_CHECK();
    std::unique_lock<std::shared_mutex> lk(x.mu);
    x.delete_locked(id);
    HANN_API_END
}

int hann_pqivf_bulk_delete(hann_pqivf *idx, const int64_t *ids, size_t n) {
    HANN_API_BEGIN
    IDX_CHECK();
    std::unique_lock<std::shared_mutex> lk(x.mu);
    x.bulk_delete(ids, n);
    HANN_API_END
}

int hann_pqivf_update(hann_pqivf *idx, int64_t id, const float *vec, size_t len) {
    HANN_API_BEGIN
    IDX_CHECK();
    std::unique_lock<std::shared_mutex> lk(x.mu);
    x.update(id, vec, len);
    HANN_API_END
}

int hann_pqivf_bulk_update(hann_pqivf *idx, const int64_t *ids, const float *flat, const size_t *lens, size_t n) {
    HANN_API_BEGIN
    IDX_CHECK();
    std::unique_lock<std::shared_mutex> lk(x.mu);
    x.bulk_update(ids, flat, lens, n);
    HANN_API_END
}

int hann_pqivf_train(hann_pqivf *idx) {
    HANN_API_BEGIN
    IDX_CHECK();
    std::unique_lock<std::shared_mutex> lk(x.mu);
    x.train();
    HANN_API_END
}

void hann_search_default_params(hann_search_params *p) {
    std::memset(p, 0, sizeof(*p));
    p->scan_mode = HANN_SCAN_EXACT;
}

int hann_pqivf_search(hann_pqivf *idx, const float *query, size_t len, int64_t k, int64_t *out_ids, double *out_dist,
                      int64_t *out_count) {
    HANN_API_BEGIN
    IDX_CHECK();
    *out_count = 0;
    std::unique_lock<std::shared_mutex> lk(x.mu);   // the device mirror may be rebuilt
    x.check_searchable(k, len);
    if (k > 1 << 20) fail(HANN_ERR_UNSUPPORTED, "k above 2^20 is not supported");
    hann_search_params p;
    hann_search_default_params(&p);
    int32_t kk = (int32_t)k, cnt = 0;
    x.search_host(query, 1, &kk, nullptr, p, out_ids, out_dist, &cnt);
    *out_count = cnt;
    HANN_API_END
}

int hann_pqivf_search_batch(hann_pqivf *idx, const float *queries, size_t nq, const int32_t *k, const int32_t *nprobe,
                            const hann_search_params *params, int64_t *out_ids, double *out_dist, int32_t *out_counts) {
    HANN_API_BEGIN
    IDX_CHECK();
    if (!k) fail(HANN_ERR_INVALID_ARG, "k array is null");
    std::unique_lock<std::shared_mutex> lk(x.mu);
    hann_search_params p;
    hann_search_default_params(&p);
    if (params) p = *params;
    x.search_host(queries, (int64_t)nq, k, nprobe, p, out_ids, out_dist, out_counts);
    HANN_API_END
}

int hann_pqivf_search_batch_dev(hann_pqivf *idx, const float *queries, size_t nq, const int32_t *k,
                                const int32_t *nprobe, int32_t kmax, int32_t max_nprobe,
                                const hann_search_params *params, int64_t *out_ids, float *out_dist,
                                int32_t *out_counts, void *stream) {
    HANN_API_BEGIN
    IDX_CHECK();
    std::unique_lock<std::shared_mutex> lk(x.mu);
    x.check_searchable(kmax, (size_t)x.dim);
    hann_search_params p;
    hann_search_default_params(&p);
    if (params) p = *params;
    x.sync_device();
    hipStream_t st = stream ? (hipStream_t)stream : x.stream;
    Workspace &w = stream ? x.stream_ws(st) : x.ws;
    float *key = w.get<float>(11, nq * (size_t)kmax);
    x.search_device(queries, (int64_t)nq, k, nprobe, kmax, max_nprobe > 0 ? max_nprobe : (int)x.nprobe, p, key, out_ids,
                    out_counts, st, -1, &w);
    launch_finalize(key, (int64_t)nq * kmax, out_dist, st);
    if (!stream)
        HANN_HIP(hipStreamSynchronize(st));
    else
        x.note_inflight(st);
    HANN_API_END
}

// A device search bound to fixed buffers, replayed as one HIP graph. The
// first launch (and the first after the index's device mirror changed) runs
// the search eagerly, which also sizes the graph's own workspace and fills
// the index's caches, then records the same launches into a graph; later
// launches replay it. Shapes whose search waits on the host (every query
// probing every list with per-query nprobe, or a certified coarse step over
// more centroids than it ranks on the device) and profiled searches run
// eagerly on every launch.
struct hann_search_graph {
    hann_pqivf *idx = nullptr;
    const float *queries = nullptr;
    size_t nq = 0;
    const int32_t *k = nullptr, *nprobe = nullptr;
    int32_t kmax = 0, max_nprobe = 0;
    hann_search_params p{};
    int64_t *out_ids = nullptr;
    float *out_dist = nullptr;
    int32_t *out_counts = nullptr;
    hipStream_t st = nullptr;
    bool sync_after = false;
    hann::Workspace ws;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t epoch = ~0ull;
    int64_t captures = 0, replays = 0, eager = 0;
    void drop() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~hann_search_graph() { drop(); }

    void search(hann::PQIVF &x) {
        float *key = ws.get<float>(11, nq * (size_t)kmax);
        x.search_device(queries, (int64_t)nq, k, nprobe, kmax, max_nprobe > 0 ? max_nprobe : (int)x.nprobe, p, key,
                        out_ids, out_counts, st, -1, &ws);
        hann::launch_finalize(key, (int64_t)nq * kmax, out_dist, st);
    }

    bool capturable(const hann::PQIVF &x) const {
        if (x.profile) return false;
        const int64_t K = (int64_t)x.lists.size();
        const int64_t np = max_nprobe > 0 ? max_nprobe : (int64_t)x.nprobe;
        if (np >= K && nprobe) return false;   // the smallest nprobe is read on the host
        if (np < K && p.coarse_precision == HANN_PREC_BF16 && K > hann::certified_no_sync_max_rows()) return false;
        return true;
    }
};

int hann_pqivf_search_graph_create(hann_pqivf *idx, const float *queries, size_t nq, const int32_t *k,
                                   const int32_t *nprobe, int32_t kmax, int32_t max_nprobe,
                                   const hann_search_params *params, int64_t *out_ids, float *out_dist,
                                   int32_t *out_counts, void *stream, hann_search_graph **out) {
    HANN_API_BEGIN
    IDX_CHECK();
    if (!out) fail(HANN_ERR_INVALID_ARG, "output handle pointer is null");
    *out = nullptr;
    if (nq == 0) fail(HANN_ERR_INVALID_ARG, "a search graph needs at least one query");
    if (!queries || !k || !out_ids || !out_dist || !out_counts)
        fail(HANN_ERR_INVALID_ARG, "queries, k and the outputs must be device pointers");
    {
        std::unique_lock<std::shared_mutex> lk(x.mu);
        x.check_searchable(kmax, (size_t)x.dim);
        x.ensure_stream();
    }
    std::unique_ptr<hann_search_graph> g(new hann_search_graph);
    g->idx = idx;
    g->queries = queries;
    g->nq = nq;
    g->k = k;
    g->nprobe = nprobe;
    g->kmax = kmax;
    g->max_nprobe = max_nprobe;
    hann_search_default_params(&g->p);
    if (params) g->p = *params;
    g->out_ids = out_ids;
    g->out_dist = out_dist;
    g->out_counts = out_counts;
    // the legacy default stream cannot be captured: the index's own stream
    // runs the graph then, and the launch waits for it like the eager call
    g->sync_after = stream == nullptr;
    g->st = stream ? (hipStream_t)stream : x.stream;
    *out = g.release();
    HANN_API_END
}

int hann_search_graph_launch(hann_search_graph *g) {
    HANN_API_BEGIN
    if (!g || !g->idx) fail(HANN_ERR_INVALID_ARG, "search graph handle is null");
    PQIVF &x = g->idx->ix;
    std::unique_lock<std::shared_mutex> lk(x.mu);
    x.check_searchable(g->kmax, (size_t)x.dim);
    x.sync_device();
    if (!g->capturable(x)) {
        g->drop();
        g->search(x);
        ++g->eager;
    } else if (g->exec && g->epoch == x.mirror_epoch) {
        HANN_HIP(hipGraphLaunch(g->exec, g->st));
        ++g->replays;
    } else {
        g->drop();
        // this launch's results come from the eager run; the capture only
        // records the launches (nothing runs while the stream captures)
        g->search(x);
        ++g->eager;
        HANN_HIP(hipStreamBeginCapture(g->st, hipStreamCaptureModeRelaxed));
        try {
            g->search(x);
        } catch (...) {
            hipGraph_t partial = nullptr;
            (void)hipStreamEndCapture(g->st, &partial);
            if (partial) (void)hipGraphDestroy(partial);
            throw;
        }
        HANN_HIP(hipStreamEndCapture(g->st, &g->graph));
        HANN_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
        g->epoch = x.mirror_epoch;
        ++g->captures;
    }
    if (g->sync_after)
        HANN_HIP(hipStreamSynchronize(g->st));
    else
        x.note_inflight(g->st);
    HANN_API_END
}

int hann_search_graph_counts(hann_search_graph *g, int64_t *captures, int64_t *replays, int64_t *eager) {
    HANN_API_BEGIN
    if (!g) fail(HANN_ERR_INVALID_ARG, "search graph handle is null");
    if (captures) *captures = g->captures;
    if (replays) *replays = g->replays;
    if (eager) *eager = g->eager;
    HANN_API_END
}

void hann_search_graph_destroy(hann_search_graph *g) {
    if (!g) return;
    // the graph's workspace may still be in use by queued work
    if (g->st) (void)hipStreamSynchronize(g->st);
    delete g;
}

int hann_pqivf_stats(hann_pqivf *idx, hann_index_stats *out) {
    HANN_API_BEGIN
    IDX_CHECK();
    std::shared_lock<std::shared_mutex> lk(x.mu);
    std::memset(out, 0, sizeof(*out));
    out->count = x.count();
    out->dimension = x.dim;
    out->fallback_searches = x.fallback_count();
    std::strncpy(out->distance, "euclidean", sizeof(out->distance) - 1);
    HANN_API_END
}

int hann_pqivf_save(hann_pqivf *idx, void **buf, size_t *len) {
    HANN_API_BEGIN
    IDX_CHECK();
    std::shared_lock<std::shared_mutex> lk(x.mu);
    std::vector<unsigned char> b = x.save();
    void *p = std::malloc(b.size() ? b.size() : 1);
    if (!p) fail(HANN_ERR_OOM, "host allocation failed");
    std::memcpy(p, b.data(), b.size());
    *buf = p;
    *len = b.size();
    HANN_API_END
}

int hann_pqivf_load(hann_pqivf *idx, const void *buf, size_t len) {
    HANN_API_BEGIN
    IDX_CHECK();
    std::unique_lock<std::shared_mutex> lk(x.mu);
    x.load(buf, len);
    HANN_API_END
}

namespace {
int give_buffer(std::vector<unsigned char> &&b, void **buf, size_t *len) {
    if (!buf || !len) fail(HANN_ERR_INVALID_ARG, "output buffer pointers are null");
    void *p = std::malloc(b.size() ? b.size() : 1);
    if (!p) fail(HANN_ERR_OOM, "host allocation failed");
    std::memcpy(p, b.data(), b.size());
    *buf = p;
    *len = b.size();
    return HANN_OK;
}
}  // namespace

int hann_pqivf_gob_encode(hann_pqivf *idx, int format_version, void **buf, size_t *len) {
    HANN_API_BEGIN
    IDX_CHECK();
    std::shared_lock<std::shared_mutex> lk(x.mu);
    give_buffer(x.gob_encode(format_version), buf, len);
    HANN_API_END
}

int hann_pqivf_gob_decode(hann_pqivf *idx, const void *buf, size_t len) {
    HANN_API_BEGIN
    IDX_CHECK();
    if (!buf && len) fail(HANN_ERR_INVALID_ARG, "buffer is null");
    std::unique_lock<std::shared_mutex> lk(x.mu);
    x.gob_decode(buf, len);
    HANN_API_END
}

int hann_pqivf_save_gob(hann_pqivf *idx, int format_version, void **buf, size_t *len) {
    HANN_API_BEGIN
    IDX_CHECK();
    std::shared_lock<std::shared_mutex> lk(x.mu);
    give_buffer(x.gob_save(format_version), buf, len);
    HANN_API_END
}

int hann_pqivf_load_gob(hann_pqivf *idx, const void *buf, size_t len) {
    HANN_API_BEGIN
    IDX_CHECK();
    if (!buf && len) fail(HANN_ERR_INVALID_ARG, "buffer is null");
    std::unique_lock<std::shared_mutex> lk(x.mu);
    x.gob_load(buf, len);
    HANN_API_END
}

int hann_pqivf_profile(hann_pqivf *idx, int enable) {
    HANN_API_BEGIN
    IDX_CHECK();
    std::unique_lock<std::shared_mutex> lk(x.mu);
    x.profile = enable != 0;
    HANN_API_END
}

int hann_pqivf_stage_times(hann_pqivf *idx, double *coarse_ms, double *scan_ms, double *rerank_ms, int64_t *calls) {
    HANN_API_BEGIN
    IDX_CHECK();
    std::unique_lock<std::shared_mutex> lk(x.mu);
    x.resolve_timers();
    *coarse_ms = x.t_coarse_ms;
    *scan_ms = x.t_scan_ms;
    *rerank_ms = x.t_rerank_ms;
    *calls = x.t_calls;
    x.t_coarse_ms = x.t_scan_ms = x.t_rerank_ms = 0;
    x.t_calls = 0;
    HANN_API_END
}

int hann_pqivf_scan_kernel_time(hann_pqivf *idx, double *kernel_ms) {
    HANN_API_BEGIN
    IDX_CHECK();
    std::unique_lock<std::shared_mutex> lk(x.mu);
    x.resolve_timers();
    *kernel_ms = x.t_kernel_ms;
    x.t_kernel_ms = 0;
    HANN_API_END
}

int hann_pqivf_scan_kernel_times(hann_pqivf *idx, double *filter_ms, double *sample_ms) {
    HANN_API_BEGIN
    IDX_CHECK();
    std::unique_lock<std::shared_mutex> lk(x.mu);
    x.resolve_timers();
    *filter_ms = x.t_kernel_ms;
    *sample_ms = x.t_sample_ms;
    x.t_kernel_ms = x.t_sample_ms = 0;
    HANN_API_END
}

int hann_pqivf_last_tune(hann_pqivf *idx, char *buf, size_t len) {
    HANN_API_BEGIN
    IDX_CHECK();
    std::unique_lock<std::shared_mutex> lk(x.mu);
    if (!buf || len == 0) fail(HANN_ERR_INVALID_ARG, "buffer is null or empty");
    const size_t n = std::min(len - 1, x.last_tune.size());
    std::memcpy(buf, x.last_tune.data(), n);
    buf[n] = 0;
    HANN_API_END
}

int hann_pqivf_last_scan(hann_pqivf *idx, int *mode, int *fused) {
    HANN_API_BEGIN
    IDX_CHECK();
    std::shared_lock<std::shared_mutex> lk(x.mu);
    *mode = x.last_scan_mode.load();
    *fused = x.last_scan_fused.load();
    HANN_API_END
}

int hann_pqivf_export_lists(hann_pqivf *idx, int64_t *list_off, int64_t *ids, int32_t *codes, float *vecs) {
    HANN_API_BEGIN
    IDX_CHECK();
    std::shared_lock<std::shared_mutex> lk(x.mu);
    int64_t pos = 0;
    for (size_t c = 0; c < x.lists.size(); ++c) {
        const ListData &L = x.lists[c];
        if (list_off) list_off[c] = pos;
        const size_t n = L.ids.size();
        if (ids) std::memcpy(ids + pos, L.ids.data(), n * sizeof(int64_t));
        if (codes)
            for (size_t r = 0; r < n; ++r)
                for (int64_t i = 0; i < x.m; ++i)
                    codes[(pos + (int64_t)r) * x.m + i] = L.has_codes[r] ? L.codes[r * x.m + i] : -1;
        if (vecs) std::memcpy(vecs + pos * x.dim, L.vecs.data(), L.vecs.size() * sizeof(float));
        pos += (int64_t)n;
    }
    if (list_off) list_off[x.lists.size()] = pos;
    HANN_API_END
}

int hann_pqivf_is_trained(hann_pqivf *idx) {
    if (!idx) return 0;
    std::shared_lock<std::shared_mutex> lk(idx->ix.mu);
    return idx->ix.trained ? 1 : 0;
}

int hann_pqivf_shape(hann_pqivf *idx, int64_t *k, int64_t *m, int64_t *p, int64_t *dim) {
    HANN_API_BEGIN
    IDX_CHECK();
    std::shared_lock<std::shared_mutex> lk(x.mu);
    *k = x.kc;
    *m = x.m;
    *p = x.pmax;
    *dim = x.dim;
    HANN_API_END
}

int hann_pqivf_coarse_centroids(hann_pqivf *idx, float *out) {
    HANN_API_BEGIN
    IDX_CHECK();
    std::shared_lock<std::shared_mutex> lk(x.mu);
    std::memcpy(out, x.coarse.data(), x.coarse.size() * sizeof(float));
    HANN_API_END
}

int hann_pqivf_codebooks(hann_pqivf *idx, float *out, int64_t *cb_len) {
    HANN_API_BEGIN
    IDX_CHECK();
    std::shared_lock<std::shared_mutex> lk(x.mu);
    std::memcpy(out, x.codebooks.data(), x.codebooks.size() * sizeof(float));
    for (int64_t i = 0; i < x.m; ++i) cb_len[i] = x.have_codebooks ? x.cb_len[i] : 0;
    HANN_API_END
}

int hann_pqivf_list_ids(hann_pqivf *idx, int64_t list, int64_t *out, int64_t *len) {
    HANN_API_BEGIN
    IDX_CHECK();
    std::shared_lock<std::shared_mutex> lk(x.mu);
    if (list < 0 || list >= (int64_t)x.lists.size() || !x.lists[list].exists) {
        *len = -1;   // no such list key
        return HANN_OK;
    }
    const ListData &L = x.lists[list];
    *len = (int64_t)L.ids.size();
    if (out) std::memcpy(out, L.ids.data(), L.ids.size() * sizeof(int64_t));
    HANN_API_END
}

int hann_pqivf_list_codes(hann_pqivf *idx, int64_t list, int32_t *out, int64_t *len) {
    HANN_API_BEGIN
    IDX_CHECK();
    std::shared_lock<std::shared_mutex> lk(x.mu);
    if (list < 0 || list >= (int64_t)x.lists.size() || !x.lists[list].exists) {
        *len = -1;
        return HANN_OK;
    }
    const ListData &L = x.lists[list];
    *len = (int64_t)L.ids.size();
    if (out)
        for (size_t r = 0; r < L.ids.size(); ++r)
            for (int64_t i = 0; i < x.m; ++i) out[r * x.m + i] = L.has_codes[r] ? L.codes[r * x.m + i] : -1;
    HANN_API_END
}

int hann_pqivf_encode_vector(hann_pqivf *idx, const float *vec, size_t len, int64_t cluster, int32_t *codes) {
    HANN_API_BEGIN
    IDX_CHECK();
    std::unique_lock<std::shared_mutex> lk(x.mu);
    // encodeVector (pqivf/index.go:669-675)
    if (!x.have_codebooks) fail(HANN_ERR_INVALID_ARG, "codebooks not trained");
    if ((int64_t)len != x.dim) fail(HANN_ERR_INVALID_ARG, "vector lengths do not match");
    if (cluster < 0 || cluster >= x.kc) fail(HANN_ERR_INVALID_ARG, "cluster out of range");
    for (int64_t i = 0; i < x.m; ++i)
        if (x.cb_len[i] == 0) fail(HANN_ERR_INVALID_ARG, "failed to encode sub-vector");
    int32_t c = (int32_t)cluster;
    x.encode_rows(vec, 1, &c, codes);
    HANN_API_END
}

}  // extern "C"

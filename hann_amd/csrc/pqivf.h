// pqivf.h -- the PQIVF index runtime (pqivf.cpp).
#pragma once

#include <array>
#include <atomic>
#include <map>
#include <mutex>
#include <set>
#include <shared_mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "batch.h"
#include "ivf.h"
#include "kmeans.h"
#include "runtime.h"
#include "scan_list.h"

namespace hann {

// Entries of one coarse list in storage order.
struct ListData {
    std::vector<int64_t> ids;
    std::vector<float> vecs;        // [len][dim]
    std::vector<int32_t> codes;     // [len][m]
    std::vector<uint8_t> has_codes; // entry carries PQ codes
    bool exists = false;            // the map key invertedLists[c] is present
};

// Vectors added before training (pendingVectors), stored contiguously:
// Train sorts them by id, so their storage order carries no meaning.
struct PendingStore {
    int64_t dim = 0;
    std::unordered_map<int64_t, int64_t> slot;
    std::vector<int64_t> ids;
    std::vector<float> vecs;   // [size][dim]
    bool count(int64_t id) const { return slot.count(id) != 0; }
    size_t size() const { return ids.size(); }
    void put(int64_t id, const float *v) {
        slot[id] = (int64_t)ids.size();
        ids.push_back(id);
        vecs.insert(vecs.end(), v, v + dim);
    }
    bool erase(int64_t id) {
        auto it = slot.find(id);
        if (it == slot.end()) return false;
        const int64_t s = it->second, last = (int64_t)ids.size() - 1;
        slot.erase(it);
        if (s != last) {
            ids[s] = ids[last];
            std::copy(vecs.begin() + last * dim, vecs.begin() + (last + 1) * dim, vecs.begin() + s * dim);
            slot[ids[s]] = s;
        }
        ids.pop_back();
        vecs.resize((size_t)last * dim);
        return true;
    }
    void clear() {
        slot.clear();
        ids.clear();
        vecs.clear();
        ids.shrink_to_fit();
        vecs.shrink_to_fit();
    }
};

// runKMeans over `groups` problems on device data; returns min(k, n)
// precision HANN_PREC_BF16 (single group, dsub >= 32): certified BF16
// assignment, identical results. assign_out (or null): [groups][n], holds
// the last Lloyd iteration's assignment (made before its update) on return
int64_t run_kmeans(const float *data, int64_t n, int64_t row_stride, int dsub, int k, int iters, int groups,
                   std::vector<uint64_t> &states, float *cents, KMeansScratch &sc, hipStream_t st,
                   int precision = HANN_PREC_FP32, int init_mode = 0, int32_t *assign_out = nullptr);
// assignPoints in either precision (BF16 needs a contiguous [n][d] layout)
void assign_points(const float *data, int64_t n, const float *cents, int k, int d, int precision,
                   const int32_t *hint, int32_t *out, hipStream_t st);

// A whole index state read from a file, installed at once.
struct IndexState {
    int64_t dim = 0, coarse_k = 0, m = 0, pq_k = 0, iters = 0, nprobe = 0;
    bool fallback = false, trained = false;
    bool has_seed = false;   // the file carries the seed and training epoch
    uint64_t seed = 0, train_epoch = 0;
    int64_t kc = 0;
    std::vector<float> coarse;
    bool have_codebooks = false;
    int64_t pmax = 0;
    std::vector<int64_t> cb_len;
    std::vector<float> codebooks;
    std::vector<ListData> lists;
    std::unordered_map<int64_t, int64_t> id_to_cluster;
    PendingStore pending;
};

struct PQIVF {
    // parameters (pqivf/index.go:36-54)
    int64_t dim = 0, coarse_k = 16, m = 1, pq_k = 16, iters = 10, nprobe = 3;
    bool fallback = true;
    uint64_t seed = 0;
    int train_precision = 0, kmeans_init = 0;
    int64_t coarse_sample = 0, pq_sample = 0;
    // state
    bool trained = false;
    int64_t kc = 0;                   // len(coarseCentroids)
    std::vector<float> coarse;        // [kc][dim]
    bool have_codebooks = false;
    int64_t pmax = 0;                 // codebook rows (stride)
    std::vector<int64_t> cb_len;      // [m]
    std::vector<float> codebooks;     // [m][pmax][dim/m]
    std::vector<ListData> lists;      // [kc]
    std::unordered_map<int64_t, int64_t> id_to_cluster;
    PendingStore pending;
    uint64_t train_epoch = 0;
    std::shared_mutex mu;

    // device mirror
    hipStream_t stream = nullptr;
    Workspace ws;
    // the device-buffer search's workspace per caller stream, so that
    // searches on different streams can run on the device at once (each
    // stream's searches reuse its buffers in stream order)
    std::map<hipStream_t, std::unique_ptr<Workspace>> ws_by_stream;
    Workspace &stream_ws(hipStream_t s) {
        std::unique_ptr<Workspace> &w = ws_by_stream[s];
        if (!w) w.reset(new Workspace);
        return *w;
    }
    // searches queued on callers' streams (the device-buffer search and
    // graph launches return before their work ends): one event per stream,
    // recorded after that stream's latest search. A sync that changes the
    // device mirror, and freeing the index, first wait for all of them, so
    // no search still running reads lists that are being moved.
    std::map<hipStream_t, hipEvent_t> inflight;
    void note_inflight(hipStream_t s);
    void wait_inflight();
    bool dirty = true;
    // bumped by every device-mirror update: captured search graphs hold
    // buffer pointers and sizes of the mirror they saw
    uint64_t mirror_epoch = 0;
    bool full_rebuild = true;
    // incremental mirror updates: lists have slack capacity on the device;
    // appended tails or whole changed lists are uploaded at the next sync
    std::vector<int64_t> dev_start, dev_cap, dev_len, append_from;
    std::vector<uint8_t> touch_flags;   // bit 0 appends, bit 1 deletes, bit 2 other changes
    std::vector<std::vector<int64_t>> deleted_ids;   // per list, since the last sync
    std::vector<int64_t> touched_lists;
    int64_t incremental_syncs = 0, full_syncs = 0;
    bool any_uncoded = false;
    float cb_fp8_scale = 1.f;   // FP8 list scan: the codebook scale (a power of two)
    int code_bytes = 1;
    int64_t n_entries = 0;
    DevBuf<int64_t> d_list_start, d_list_len, d_ids, d_cb_len, all_lists;
    DevBuf<uint8_t> d_codes, d_has_codes;
    DevBuf<float> d_vecs, d_coarse, d_cb, scratch_f;
    DevBuf<unsigned long long> d_fb_counter;
    Workspace mws;   // grow-only device scratch of the mutation paths (no allocation per call)
    // persistent scan: per-entry list terms, kept current by every sync
    DevBuf<float> d_term;
    DevBuf<int64_t> d_term_lists;
    DevBuf<unsigned> d_ymax_tmax;   // max |y_e|, max |t_e| (float bits) over the entries
    // the scan the last search ran (searches may run concurrently under
    // the shared lock, so these are atomics)
    std::atomic<int> last_scan_mode{-1}, last_scan_fused{0};
    void refresh_terms(const std::vector<int64_t> *lists);   // null: every list
    // the coarse step's BF16 candidate pass (used under the index lock) and
    // the centroids' BF16 copy (rebuilt after a full sync)
    Context coarse_ctx;
    BF16Base coarse_bf16;
    std::map<uint64_t, int> tuned_wpq;   // persistent scan waves per query; key: (cap, maxp bucket)
    // list-major scans' work items and partitions; key: (scan, selection
    // size, nprobe bucket). Searches hold the index lock exclusively, so the
    // maps need no lock of their own.
    std::map<uint64_t, ListTune> tuned_list;
    int num_cus = 0;
    std::string last_tune;

    PQIVF();
    ~PQIVF();
    void ensure_stream();
    void validate_vector(size_t len) const;
    bool codes_valid(const int32_t *codes) const;
    int64_t nearest_centroid(const float *vec);
    void encode_rows(const float *vecs, int64_t n, const int32_t *clusters, int32_t *codes_out);
    void assign_locked(const std::vector<int64_t> &ids, const std::vector<const float *> &vecs);
    void add_locked(int64_t id, const float *vec, size_t len);
    void bulk_add(const int64_t *ids, const float *flat, const size_t *lens, size_t n);
    void remove_from_list(int64_t id, int64_t cluster);
    void delete_locked(int64_t id);
    void bulk_delete(const int64_t *ids, size_t n);
    void update(int64_t id, const float *vec, size_t len);
    void bulk_update(const int64_t *ids, const float *flat, const size_t *lens, size_t n);
    void train();
    void reserve_slack(ListData &L, int64_t len) const;
    void mark_dirty();           // the next sync rebuilds the whole device mirror
    void touch_append(int64_t l);   // entries are about to be appended to list l
    void touch_other(int64_t l);    // list l changed in any other way
    void touch_delete(int64_t l, int64_t id);   // id was removed from list l
    void sync_device();
    void check_searchable(int64_t k, size_t len) const;
    void search_device(const float *dq, int64_t nq, const int32_t *dk, const int32_t *dnp, int kmax, int pmax_probe,
                       const hann_search_params &p, float *out_key, int64_t *out_id, int32_t *out_cnt, hipStream_t st,
                       int min_nprobe = -1, Workspace *wsp = nullptr);
    int device_min(const int32_t *v, int64_t n, hipStream_t st, Workspace &w);
    void search_host(const float *queries, int64_t nq, const int32_t *k, const int32_t *np,
                     const hann_search_params &p, int64_t *out_ids, double *out_dist, int32_t *out_counts);
    int64_t fallback_count();
    void run_fused_scan(const ScanArgs &a, int64_t nq, int maxp, const double *probe_rank, hipStream_t st,
                        Workspace &w, bool main_scan = true);
    // the list-major scan's batch buffers for nq queries probing np lists
    // with rerank depth selk: pairs, work items, sample-row slots, record
    // room per query, and the arena's bytes
    struct ListPlan {
        int64_t pairs_max = 0, items_max = 0, sample_slots = 0, capq = 0;
        size_t arena = 0;
        bool exact = false;   // the exact scan's work items
    };
    ListPlan list_plan(int64_t nq, int np, int selk, bool exact = false) const;
    // the device lists' longest length and chunk count, refreshed from the
    // mirror when it changed (callers hold the index lock); readable without
    // it, then as of the last refresh
    mutable std::atomic<uint64_t> shape_epoch{~0ull};
    // Σ ceil(len / 2^c) over the device lists for c = 6 .. 11
    mutable std::atomic<int64_t> shape_max_len{0}, shape_chunks[6]{};
    void refresh_shape() const;
    // reasons the list-major scan was declined, each said once on stderr
    std::mutex note_mu;
    std::set<std::string> list_notes;
    // an upper bound of the device workspace bytes search_device takes for
    // nq queries, k up to kmax and nprobe up to pmax_probe (every slot of a
    // grow-only workspace counted at the size this shape asks of it; list
    // lengths as of the last refresh_shape)
    int64_t search_scratch_bytes(int64_t nq, int kmax, int pmax_probe, const hann_search_params &p) const;
    static bool list_plan_fits(const ListPlan &L, int64_t nq);
    // list-major scan (x3: 0 BF16, 1 split BF16 MFMA, 2 exact FP32); a.out_*
    // are the candidate rows of the rerank
    void run_list_scan(const ScanArgs &a, int64_t nq, int x3, int maxp, const double *probe_rank, hipStream_t st,
                       Workspace &w, const ListPlan &L, const ListTune &t);
    // the configuration of the list-major scan for this batch shape: tuned
    // on its first batch (every candidate timed once on the search stream,
    // the fastest kept; results do not depend on it), the default while the
    // stream is being captured
    ListTune list_tuned(const ScanArgs &a, int64_t nq, int x3, int maxp, const double *probe_rank, hipStream_t st,
                        Workspace &w, const ListPlan &L);
    // stage timers (HIP events on the search stream), accumulated when enabled
    bool profile = false;
    double t_coarse_ms = 0, t_scan_ms = 0, t_rerank_ms = 0;
    int64_t t_calls = 0;
    // per search: stage boundaries [0..3]; [4], [5] bracket the main scan
    // kernel's launch (k_scan_list's filter pass, k_scan_block or k_scan)
    // and [6], [7] the list-major scan's sample pass, when recorded
    std::vector<std::array<hipEvent_t, 8>> t_events;   // pending, resolved by stage_times()
    hipEvent_t *kmark = nullptr;   // the current search's [4] .. [7] while it launches
    double t_kernel_ms = 0, t_sample_ms = 0;
    int64_t t_kernel_calls = 0;
    void mark_kernel(int which, hipStream_t st) {
        if (kmark) HANN_HIP(hipEventRecord(kmark[which], st));
    }
    void resolve_timers();
    int64_t count() const;
    std::vector<unsigned char> save() const;
    void load(const void *buf, size_t len);
    void install(IndexState &&st);
    // the reference's gob format (gob.cpp): GobEncode / GobDecode and the
    // Save / Load wrapping (pqivf/index.go:968-1101)
    std::vector<unsigned char> gob_encode(int format_version) const;
    void gob_decode(const void *buf, size_t len);
    std::vector<unsigned char> gob_save(int format_version) const;
    void gob_load(const void *buf, size_t len);
};

}  // namespace hann

// the C-ABI handle (include/hann_amd.h)
struct hann_pqivf {
    hann::PQIVF ix;
};

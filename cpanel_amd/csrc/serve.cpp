// Continuous-batching request server over one GPU (cp_server_* in cpgpu.h).
//
// A request is (geometry key + mesh, 1..64 freestream conditions, options).
// Worker threads (one per slot) take requests in arrival order; each
// geometry lives on the device once (FP32 matrices, near-field correction,
// block-Jacobi inverses, post tables), keyed by the caller's key, and has its
// own HIP stream, so requests on different geometries run on the GPU at the
// same time while the host side of new geometries (topology, CHTLS weights,
// block inversion) proceeds on other workers.  Requests for a geometry that
// is busy wait for it; when it frees up, every queued request for it with
// the same options is merged into one batched solve (up to 64 columns; each
// column's result does not depend on what else is in its batch).
//
// Admission: a new geometry is built only when its estimated device bytes
// fit the context's HBM budget; idle geometries are evicted least recently
// used first to make room.  A request that does not fit waits in the queue
// while later requests whose geometries are resident go ahead.
//
// Replaces the reference's sequential case loop (CPanelMgr.cpp:49-88, one
// cpCase::run per condition on one geometry) for many geometries and
// conditions arriving over time.
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/cpgpu.h"
#include "batched.h"
#include "device.h"
#include "precond.h"

using namespace cpgpu;

namespace cpgpu {
namespace capi {
// context lifetime (capi.cpp): the server holds one reference
void ctx_ref(cp_ctx* c);
void ctx_unref(cp_ctx* c);
// its solvers are used from several host threads from now on
void mark_served(cp_ctx* c);
}  // namespace capi
}  // namespace cpgpu

namespace {

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Req {
    uint64_t id = 0, key = 0;
    std::vector<double> xyz;
    std::vector<int32_t> tri, ids;
    std::vector<cp_freestream> fs;
    cp_solve_opts opts{};
    bool want_mu = false;
    double t_submit = 0, t_start = 0, t_ready = 0, t_end = 0;
    cp_status status = CP_OK;
    bool done = false;
    bool cache_hit = false;
    int batch_cols = 0;
    int group = 1;  // geometries of the grouped solve it ran in
    int n_body = 0;
    std::vector<cp_solve_info> info;
    std::vector<double> mu;
    std::string err;
};

struct Entry {
    cp_geom* g = nullptr;
    bool busy = false;
    bool ready = false;
    int n_body = 0;
    uint64_t est = 0;  // estimated device bytes
    double last_used = 0;
};

bool same_opts(const cp_solve_opts& a, const cp_solve_opts& b) { return std::memcmp(&a, &b, sizeof(a)) == 0; }

// device bytes of a geometry with n body panels on the mixed path
uint64_t estimate_bytes(const Req& r) {
    const int64_t nt = (int64_t)r.ids.size();
    int64_t nb = 0, nw = 0;
    for (int32_t id : r.ids) (id > 1000 ? nw : nb)++;
    (void)nt;
    const int64_t lda = round_up(nb, 512);
    uint64_t b = 3ull * lda * (nb + 15) * 4;               // A32, B32, bf16 tiles of A32
    b += (uint64_t)lda * std::max<int64_t>(nw, 1) * 8;     // wake block W
    b += (uint64_t)nb * 120 * 28;                          // near-field CSR (~120 entries per row)
    // block-Jacobi: FP32 inverses + their packed copy + the FP64 inversion
    // workspace, blocks of about bj_block_target(nb) rows
    b += (uint64_t)nb * std::min<int64_t>(nb, bj_block_target((int)nb)) * (4 + 4 + 8);
    b += (uint64_t)nb * (16 * 4 + 90 * 8 + 64);            // post tables
    b += (uint64_t)BatchedSolver::workspace_bytes(lda);    // solver workspace
    b += (uint64_t)lda * 64 * 8 * 3;                       // per-call temporaries
    return b + (64ull << 20);
}

}  // namespace

struct cp_server {
    cp_ctx* ctx = nullptr;
    int slots = 1;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req*> queue;
    std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
    std::unordered_map<uint64_t, Entry> cache;
    std::vector<std::thread> workers;
    bool stop = false;
    uint64_t next_id = 1;
    cp_server_stats st{};
    uint64_t budget = 0;
    // grouping of small resident geometries (see cp_server_create's header
    // comment): geometries per group and their panel limit
    int group_max = 8;
    int group_n = 4096;

    // device bytes of resident geometries (estimates) + entries being built
    uint64_t resident() const {
        uint64_t s = 0;
        for (auto& kv : cache) s += kv.second.est;
        return s;
    }

    // evicts idle geometries (LRU first) until `need` more bytes fit; true if
    // they fit.  Called with mu held.
    bool make_room(uint64_t need) {
        if (!budget) return true;
        while (resident() + need > budget) {
            uint64_t victim = 0;
            double oldest = 1e300;
            bool found = false;
            for (auto& kv : cache)
                if (!kv.second.busy && kv.second.ready && kv.second.last_used < oldest) {
                    oldest = kv.second.last_used;
                    victim = kv.first;
                    found = true;
                }
            if (!found) return false;
            cp_geometry_release(cache[victim].g);
            cache.erase(victim);
            st.evictions++;
        }
        return true;
    }

    // r and every later queued request for its geometry with its options,
    // oldest first, up to 64 columns, taken off the queue (mu held)
    std::vector<Req*> take_batch(Req* r) {
        std::vector<Req*> batch;
        int cols = 0;
        for (auto it = queue.begin(); it != queue.end();) {
            Req* q = *it;
            if (q->key == r->key && same_opts(q->opts, r->opts) &&
                (batch.empty() || cols + (int)q->fs.size() <= kRhsTile)) {
                batch.push_back(q);
                cols += (int)q->fs.size();
                it = queue.erase(it);
            } else {
                ++it;
            }
        }
        return batch;
    }

    // next runnable request (mu held): the oldest whose geometry is idle and
    // resident, or absent and admissible
    Req* pick() {
        for (Req* r : queue) {
            auto it = cache.find(r->key);
            if (it != cache.end()) {
                if (!it->second.busy) return r;
                continue;
            }
            uint64_t need = estimate_bytes(*r);
            if (make_room(need)) return r;
        }
        return nullptr;
    }

    void worker() {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            Req* r = nullptr;
            while (!stop && !(r = pick())) cv_work.wait(lk);
            if (stop) return;
            // batch: this request and every queued one for the same geometry
            // and options, oldest first, up to 64 columns
            std::vector<Req*> batch = take_batch(r);
            Entry& e = cache[r->key];
            const bool hit = e.ready;
            if (!hit) e.est = estimate_bytes(*r);
            e.busy = true;
            // other idle resident small geometries with queued requests of the
            // same options join a small resident geometry's solve as a group
            std::vector<std::vector<Req*>> others;
            if (hit && group_max > 1 && e.n_body <= group_n) {
                std::vector<uint64_t> keys;
                for (Req* q : queue) {
                    if ((int)keys.size() + 1 >= group_max) break;
                    if (q->key == r->key) continue;
                    if (std::find(keys.begin(), keys.end(), q->key) != keys.end()) continue;
                    auto ce = cache.find(q->key);
                    if (ce == cache.end() || !ce->second.ready || ce->second.busy || ce->second.n_body > group_n)
                        continue;
                    keys.push_back(q->key);
                }
                for (uint64_t key : keys) {
                    Req* first = nullptr;  // the oldest queued request of that geometry
                    for (Req* q : queue)
                        if (q->key == key) {
                            first = q;
                            break;
                        }
                    others.push_back(take_batch(first));
                    cache[key].busy = true;
                }
            }
            const double t0 = now_s();
            const int ngroup = 1 + (int)others.size();
            auto start = [&](std::vector<Req*>& b, bool h) {
                int cols = 0;
                for (Req* q : b) cols += (int)q->fs.size();
                for (Req* q : b) {
                    q->t_start = t0;
                    q->cache_hit = h;
                    q->batch_cols = cols;
                    q->group = ngroup;
                }
                st.batches++;
                if (h)
                    st.cache_hits += b.size();
                else
                    st.cache_misses++;
            };
            start(batch, hit);
            for (auto& b : others) start(b, true);
            if (ngroup > 1) st.grouped_solves++;
            lk.unlock();

            cp_status s = CP_OK;
            std::string err;
            cp_geom* g = e.g;
            if (!hit) {
                cp_mesh m{(int32_t)(r->xyz.size() / 3), r->xyz.data(), (int32_t)r->ids.size(), r->tri.data(),
                          r->ids.data()};
                s = cp_geometry_upload(ctx, &m, r->key, &g);
                if (s == CP_OK) s = cp_assemble(g, CP_FP32_REFINE);
                if (s != CP_OK) err = cp_last_error(ctx);
            }
            const double t1 = now_s();
            // every batch of the solve: its geometry, conditions and results
            struct Part {
                std::vector<Req*>* b;
                cp_geom* g;
                cp_geom_info gi;
                std::vector<cp_freestream> fs;
                std::vector<cp_solve_info> info;
                std::vector<double> mu;
                bool want_mu = false;
            };
            std::vector<Part> parts(ngroup);
            {
                std::lock_guard<std::mutex> lk2(mu);  // the cache map is read under the lock
                parts[0].b = &batch;
                parts[0].g = g;
                for (int i = 1; i < ngroup; i++) {
                    parts[i].b = &others[i - 1];
                    parts[i].g = cache[others[i - 1][0]->key].g;
                }
            }
            for (Part& P : parts) {
                for (Req* q : *P.b) {
                    P.fs.insert(P.fs.end(), q->fs.begin(), q->fs.end());
                    P.want_mu = P.want_mu || q->want_mu;
                }
                P.info.resize(P.fs.size());
                P.gi = cp_geom_info{};
            }
            if (s == CP_OK) {
                for (Part& P : parts) {
                    cp_geometry_info(P.g, &P.gi);
                    if (P.want_mu) P.mu.resize((size_t)P.gi.n_body * P.fs.size());
                }
                if (ngroup == 1) {
                    Part& P = parts[0];
                    s = cp_solve(P.g, CP_FP32_REFINE, (int32_t)P.fs.size(), P.fs.data(), &r->opts,
                                 P.want_mu ? P.mu.data() : nullptr, nullptr, P.info.data());
                } else {
                    std::vector<cp_geom*> gs;
                    std::vector<int32_t> ks;
                    std::vector<const cp_freestream*> fsp;
                    std::vector<double*> mup;
                    std::vector<cp_solve_info*> infp;
                    std::vector<cp_solve_opts> op;
                    for (Part& P : parts) {
                        op.push_back((*P.b)[0]->opts);
                        gs.push_back(P.g);
                        ks.push_back((int32_t)P.fs.size());
                        fsp.push_back(P.fs.data());
                        mup.push_back(P.want_mu ? P.mu.data() : nullptr);
                        infp.push_back(P.info.data());
                    }
                    s = cp_solve_group(ngroup, gs.data(), CP_FP32_REFINE, ks.data(), fsp.data(), op.data(),
                                       mup.data(), nullptr, infp.data());
                }
                if (s != CP_OK && s != CP_NOT_CONVERGED) err = cp_last_error(ctx);
            }
            const double t2 = now_s();

            lk.lock();
            if (!hit) {
                if (g && (s == CP_OK || s == CP_NOT_CONVERGED)) {
                    e.g = g;
                    e.ready = true;
                    e.n_body = parts[0].gi.n_body;
                } else {
                    if (g) cp_geometry_release(g);
                    cache.erase(r->key);
                }
            }
            for (Part& P : parts) {
                auto it = cache.find((*P.b)[0]->key);
                if (it != cache.end()) {
                    it->second.busy = false;
                    it->second.last_used = t2;
                }
                size_t off = 0;
                for (Req* q : *P.b) {
                    const size_t k = q->fs.size();
                    q->info.assign(P.info.begin() + off, P.info.begin() + off + k);
                    if (q->want_mu && !P.mu.empty())
                        q->mu.assign(P.mu.begin() + off * P.gi.n_body, P.mu.begin() + (off + k) * P.gi.n_body);
                    q->n_body = P.gi.n_body;
                    off += k;
                    q->status = s;
                    q->err = err;
                    q->t_ready = t1;
                    q->t_end = t2;
                    q->done = true;
                    st.completed++;
                    st.conditions += k;
                }
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

extern "C" {

cp_status cp_server_create(cp_ctx* ctx, int32_t slots, uint64_t hbm_budget, cp_server** out) {
    if (!ctx || !out || slots <= 0 || slots > 16) return CP_BAD_ARG;
    *out = nullptr;
    cp_server* S = new cp_server();
    S->ctx = ctx;
    capi::ctx_ref(ctx);  // released by cp_server_destroy
    capi::mark_served(ctx);
    S->slots = slots;
    S->budget = hbm_budget;
    if (const char* e = std::getenv("CPGPU_SERVE_GROUP")) S->group_max = std::max(1, std::min(kMaxGroup, std::atoi(e)));
    if (const char* e = std::getenv("CPGPU_SERVE_GROUP_N")) S->group_n = std::max(0, std::atoi(e));
    for (int i = 0; i < slots; i++) S->workers.emplace_back([S] { S->worker(); });
    *out = S;
    return CP_OK;
}

cp_status cp_server_submit(cp_server* S, uint64_t key, const cp_mesh* mesh, int32_t k, const cp_freestream* fs,
                           const cp_solve_opts* opts, int32_t want_mu, uint64_t* id) {
    if (!S || !mesh || !fs || !opts || !id || k <= 0 || k > kRhsTile) return CP_BAD_ARG;
    if (!mesh->xyz || !mesh->tri || !mesh->ids || mesh->n_nodes <= 0 || mesh->n_tris <= 0) return CP_BAD_ARG;
    std::unique_ptr<Req> r(new Req());
    r->key = key;
    r->xyz.assign(mesh->xyz, mesh->xyz + 3 * (size_t)mesh->n_nodes);
    r->tri.assign(mesh->tri, mesh->tri + 3 * (size_t)mesh->n_tris);
    r->ids.assign(mesh->ids, mesh->ids + mesh->n_tris);
    r->fs.assign(fs, fs + k);
    r->opts = *opts;
    r->want_mu = want_mu != 0;
    r->t_submit = now_s();
    std::lock_guard<std::mutex> lk(S->mu);
    r->id = S->next_id++;
    *id = r->id;
    S->queue.push_back(r.get());
    S->reqs[r->id] = std::move(r);
    S->st.submitted++;
    S->cv_work.notify_one();
    return CP_OK;
}

cp_status cp_server_wait(cp_server* S, uint64_t id, cp_solve_info* info, double* mu, cp_request_stats* rs) {
    if (!S) return CP_BAD_ARG;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return CP_BAD_ARG;
    Req* r = it->second.get();
    S->cv_done.wait(lk, [&] { return r->done; });
    if (info) std::copy(r->info.begin(), r->info.end(), info);
    if (mu && !r->mu.empty()) std::copy(r->mu.begin(), r->mu.end(), mu);
    if (rs) {
        rs->queue_s = r->t_start - r->t_submit;
        rs->prepare_s = r->t_ready - r->t_start;
        rs->solve_s = r->t_end - r->t_ready;
        rs->latency_s = r->t_end - r->t_submit;
        rs->cache_hit = r->cache_hit;
        rs->batch_conditions = r->batch_cols;
        rs->n_body = r->n_body;
        rs->group_geometries = r->group;
    }
    cp_status s = r->status;
    S->reqs.erase(it);
    return s;
}

cp_status cp_server_get_stats(cp_server* S, cp_server_stats* out) {
    if (!S || !out) return CP_BAD_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    *out = S->st;
    out->resident_geometries = (int64_t)S->cache.size();
    out->resident_bytes_est = S->resident();
    out->device_bytes = cp_bytes_in_use(S->ctx);
    return CP_OK;
}

void cp_server_destroy(cp_server* S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    for (auto& t : S->workers) t.join();
    for (auto& kv : S->cache)
        if (kv.second.g) cp_geometry_release(kv.second.g);
    cp_ctx* c = S->ctx;
    delete S;
    capi::ctx_unref(c);
}

}  // extern "C"

// Batched and serving forms over the device pipeline (host code).
//
// Batched form: `slots` pipelines, each with its own stream, forked from and
// joined back to one origin stream.  Pairs are spread over the slots in
// longest-processing-time order of W*H*D and the kernels of different slots
// overlap on the device.  Within a slot, the pairs of the fused column-range
// schedule run as group launches: step by step (aux, Cost + Cross-H,
// right-view WTA, Cross-V, SGM, WTA, fill + median), one launch per kernel
// instance spans all of a chunk's pairs that use it (group.hpp), each pair
// on a workspace of its own; the other pairs run one at a time on the slot's
// context.  A repeated pair list replays as one hipGraph launch.  Launch
// shapes come per size bucket from an autotuner with a JSON cache.
//
// Serving form: a FIFO of host-memory requests admitted into free slots as
// long as the slot workspaces stay within an HBM budget; each admitted
// request is upload -> pipeline -> download on its slot's stream, retired
// when its completion event has fired.
//
// Every pair runs the same apg_hip_pipeline as alone: the outputs do not
// depend on the slot, the neighbours or the launch shapes.
#include <string.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "copy_pool.hpp"
#include "kernels.hpp"

namespace {

constexpr int kMaxSlots = 16;

struct BucketKey {
    int wb, hb, D, paths, cross, census;
    bool operator<(const BucketKey& o) const {
        return std::tie(wb, hb, D, paths, cross, census) < std::tie(o.wb, o.hb, o.D, o.paths, o.cross, o.census);
    }
};

BucketKey bucket_of(const apg_hip_params& p, int w, int h) {
    return {(w + 255) / 256, (h + 127) / 128, p.max_disparity - p.min_disparity, p.paths == 8 ? 8 : 4,
            p.agg_enable ? 1 : 0, p.census_type};
}

std::string device_arch(int device) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return "unknown";
    std::string a = prop.gcnArchName;
    const size_t c = a.find(':');
    return c == std::string::npos ? a : a.substr(0, c);
}

// ---------------------------------------------------------- tuning cache ----

const char* const kTuneFields[] = {"w_bucket", "h_bucket", "D", "paths", "cross", "census",
                                   "cost_segment", "sgm_dual", "cross_v_waves"};

bool json_int(const std::string& obj, const char* name, long* out) {
    const std::string key = std::string("\"") + name + "\"";
    size_t k = obj.find(key);
    if (k == std::string::npos) return false;
    k = obj.find(':', k + key.size());
    if (k == std::string::npos) return false;
    char* end = nullptr;
    *out = strtol(obj.c_str() + k + 1, &end, 10);
    return end != obj.c_str() + k + 1;
}

// Entries of a cache file written by save_cache for the same device arch.
void load_cache(const std::string& path, const std::string& arch, std::map<BucketKey, apg_hip_tuning>& table) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return;
    std::string s;
    char tmp[4096];
    size_t r;
    while ((r = fread(tmp, 1, sizeof tmp, f)) > 0) s.append(tmp, r);
    fclose(f);
    if (s.find("\"arch\": \"" + arch + "\"") == std::string::npos) return;
    size_t pos = s.find("\"entries\"");
    if (pos == std::string::npos) return;
    while ((pos = s.find('{', pos)) != std::string::npos) {
        const size_t end = s.find('}', pos);
        if (end == std::string::npos) break;
        const std::string obj = s.substr(pos, end - pos + 1);
        long v[9];
        bool ok = true;
        for (int i = 0; i < 9 && ok; ++i) ok = json_int(obj, kTuneFields[i], &v[i]);
        if (ok) {
            const BucketKey k{(int)v[0], (int)v[1], (int)v[2], (int)v[3], (int)v[4], (int)v[5]};
            table[k] = apg_hip_tuning{(int)v[6], (int)v[7], (int)v[8]};
        }
        pos = end + 1;
    }
}

bool save_cache(const std::string& path, const std::string& arch, const std::map<BucketKey, apg_hip_tuning>& table) {
    const std::string tmp = path + ".tmp";
    FILE* f = fopen(tmp.c_str(), "wb");
    if (!f) return false;
    fprintf(f, "{\n  \"format\": \"apg-sgm-amd launch shapes\",\n  \"version\": 1,\n  \"arch\": \"%s\",\n",
            arch.c_str());
    fprintf(f, "  \"entries\": [");
    bool first = true;
    for (const auto& e : table) {
        const BucketKey& k = e.first;
        const apg_hip_tuning& t = e.second;
        fprintf(f,
                "%s\n    {\"w_bucket\": %d, \"h_bucket\": %d, \"D\": %d, \"paths\": %d, \"cross\": %d, "
                "\"census\": %d, \"cost_segment\": %d, \"sgm_dual\": %d, \"cross_v_waves\": %d}",
                first ? "" : ",", k.wb, k.hb, k.D, k.paths, k.cross, k.census, t.cost_segment, t.sgm_dual,
                t.cross_v_waves);
        first = false;
    }
    fprintf(f, "\n  ]\n}\n");
    const bool ok = fclose(f) == 0;
    return ok && rename(tmp.c_str(), path.c_str()) == 0;
}

bool same_pairs(const std::vector<apg_hip_pair>& a, const apg_hip_pair* b, int n) {
    return (int)a.size() == n && (n == 0 || memcmp(a.data(), b, sizeof(apg_hip_pair) * (size_t)n) == 0);
}

int run_pair(apg_hip_ctx ctx, const apg_hip_pair& q) {
    return apg_hip_pipeline(ctx, &q.params, q.d_left, q.d_right, q.width, q.height, q.stride, q.channels,
                            q.d_disparity, q.d_confidence, q.d_invalid_reason, q.d_reliable, nullptr);
}

}  // namespace

// ================================================================= batch ====

namespace {

// Workspace bytes of the pairs one slot runs as one chunk of group launches
// (default).  Smaller chunks keep the slots at different steps, so that
// latency-bound steps (SGM) of one slot overlap issue-bound ones (Cost) of
// another: 64 uneven pairs over 4 slots took 41.3 ms in 8 GiB chunks, 37.5 ms
// in 1 GiB chunks; 64 pairs of 640x480 D=64 9.99 against 10.40 ms with each
// pair its own kernels.
constexpr size_t kGroupChunkBytes = size_t(1) << 30;
// Largest pair (W*H*D) that joins group launches (default): a larger pair's
// own grids fill the device, and in a group launch it would only pay the
// group prologue.  64 uneven pairs over 4 slots: 36.7 ms with 32 M, 37.3 with
// 48 M, 37.5 with every pair grouped, 37.3 with none.
constexpr size_t kGroupMaxStates = size_t(32) << 20;

// One group launch: its kernel instance and the offset of its chunk records
// in the plan's device buffer.
struct GroupLaunch {
    const void* kernel;
    int threads;
    size_t shm;
    uint32_t nblocks;
    size_t recs_off;
};

// The group launches of a pair list, per slot in stream order, and the
// pairs each slot runs one at a time after them.
struct GroupPlan {
    bool valid = false;
    std::vector<apg_hip_pair> pairs;
    std::vector<std::vector<GroupLaunch>> launches;
    std::vector<std::vector<int>> single;
    void* d_tables = nullptr;
    int grouped = 0, nlaunch = 0, nsingle = 0;
};

}  // namespace

struct apg_hip_batch_s {
    int device = 0;
    std::string arch;
    hipStream_t origin = nullptr;
    std::vector<apg_hip_ctx> slots;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    std::string last_error;
    std::map<BucketKey, apg_hip_tuning> tuned;
    std::vector<apg_hip_pair> graph_pairs;   // pair list of the cached graph
    bool graph_grouped = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // group launches: per slot, one context (workspace) per chunk position
    std::vector<std::vector<apg_hip_ctx>> gctx;
    GroupPlan plan;
    size_t group_bytes = kGroupChunkBytes;
    size_t group_states = kGroupMaxStates;
};

namespace {

int bfail(apg_hip_batch b, int code, const std::string& msg) {
    if (b) b->last_error = msg;
    return code;
}

int bhip(apg_hip_batch b, hipError_t e, const char* where) {
    if (e == hipSuccess) return 0;
    const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? 7 : 8;
    return bfail(b, code, std::string(where) + ": " + hipGetErrorString(e));
}

void drop_graph(apg_hip_batch b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_pairs.clear();
}

// The plan's tables are read by its launches and by a graph captured from
// them: both go.
void drop_plan(apg_hip_batch b) {
    drop_graph(b);
    if (b->plan.d_tables) {
        (void)hipStreamSynchronize(b->origin);
        for (apg_hip_ctx c : b->slots) (void)apg_hip_synchronize(c);
        (void)hipFree(b->plan.d_tables);
    }
    b->plan = GroupPlan();
}

// Longest-processing-time assignment: pairs in decreasing W*H*(D+16) (the
// +16 stands for the per-pixel work and launch latency that do not scale
// with D), each to the slot with the least work so far.
std::vector<std::vector<int>> plan_slots(const apg_hip_pair* pairs, int n, int nslots) {
    std::vector<double> wgt(n);
    for (int i = 0; i < n; ++i) {
        const apg_hip_pair& q = pairs[i];
        wgt[i] = (double)q.width * q.height * (q.params.max_disparity - q.params.min_disparity + 16);
    }
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return wgt[a] > wgt[c]; });
    std::vector<std::vector<int>> per(nslots);
    std::vector<double> load(nslots, 0.0);
    for (int i : order) {
        const int s = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        per[s].push_back(i);
        load[s] += wgt[i];
    }
    return per;
}

// Enqueues every pair on the slot streams between a fork and a join on the
// origin stream.  Slots that were forked are always joined, also on error.
int enqueue(apg_hip_batch b, const apg_hip_pair* pairs, int n) {
    const auto per = plan_slots(pairs, n, (int)b->slots.size());
    if (int s = bhip(b, hipEventRecord(b->fork, b->origin), "fork record")) return s;
    int status = 0;
    for (size_t s = 0; s < b->slots.size(); ++s) {
        if (per[s].empty()) continue;
        apg_hip_ctx ctx = b->slots[s];
        hipStream_t st = (hipStream_t)apg_hip_stream(ctx);
        if (int e = bhip(b, hipStreamWaitEvent(st, b->fork, 0), "fork wait")) return e;
        for (int i : per[s]) {
            if (status) break;
            const apg_hip_pair& q = pairs[i];
            const auto it = b->tuned.find(bucket_of(q.params, q.width, q.height));
            (void)apg_hip_set_tuning(ctx, it == b->tuned.end() ? nullptr : &it->second);
            status = run_pair(ctx, q);
            if (status) b->last_error = "pair " + std::to_string(i) + ": " + apg_hip_last_error(ctx);
        }
        hipError_t e = hipEventRecord(b->join[s], st);
        if (e == hipSuccess) e = hipStreamWaitEvent(b->origin, b->join[s], 0);
        if (e != hipSuccess && !status) status = bhip(b, e, "join");
    }
    return status;
}

void set_pair_tuning(apg_hip_batch b, apg_hip_ctx ctx, const apg_hip_pair& q) {
    const auto it = b->tuned.find(bucket_of(q.params, q.width, q.height));
    (void)apg_hip_set_tuning(ctx, it == b->tuned.end() ? nullptr : &it->second);
}

// Host image of the plan's device tables.
struct TableImage {
    std::vector<uint8_t> bytes;
    size_t put(const void* src, size_t n) {
        const size_t off = (bytes.size() + 255) & ~size_t(255);
        bytes.resize(off + n);
        memcpy(bytes.data() + off, src, n);
        return off;
    }
};

// The group launches of one step of a chunk: records with the same kernel
// instance form one launch (split where its grid would pass 2^31 threads).
int group_step(apg_hip_batch b, const std::vector<apgk::GroupRec>& recs, TableImage& img,
               std::vector<GroupLaunch>& out) {
    std::vector<bool> used(recs.size(), false);
    for (size_t i = 0; i < recs.size(); ++i) {
        if (used[i]) continue;
        const apgk::GroupRec& r0 = recs[i];
        std::vector<size_t> members;
        for (size_t j = i; j < recs.size(); ++j)
            if (!used[j] && recs[j].kernel == r0.kernel && recs[j].threads == r0.threads) {
                if (recs[j].args.size() != r0.args.size() || recs[j].rec_bytes != r0.rec_bytes)
                    return bfail(b, 8, "group records of one kernel differ");
                members.push_back(j);
                used[j] = true;
            }
        size_t m = 0;
        while (m < members.size()) {
            // one record per chunk of kGroupChunk blocks: its place in the
            // pair's grid, that grid's shape and the pair's arguments
            std::vector<uint8_t> recbuf;
            size_t shm = 0;
            uint64_t nblk = 0;
            for (; m < members.size(); ++m) {
                const apgk::GroupRec& r = recs[members[m]];
                const uint64_t count = (uint64_t)r.grid.x * r.grid.y * r.grid.z;
                const uint64_t nch = (count + apgk::kGroupChunk - 1) >> apgk::kGroupChunkLog2;
                if (count == 0) continue;
                const uint64_t next = nblk + (nch << apgk::kGroupChunkLog2);
                if (nblk && next * (uint64_t)r0.threads >= (1ull << 31)) break;
                if (next * (uint64_t)r0.threads >= (1ull << 31)) return bfail(b, 4, "pair grid too large");
                for (uint64_t k = 0; k < nch; ++k) {
                    const size_t at = recbuf.size();
                    recbuf.resize(at + r.rec_bytes, 0);
                    const apgk::GroupChunk c{(uint32_t)(k << apgk::kGroupChunkLog2), (uint32_t)count, r.grid.x,
                                             r.grid.x * r.grid.y};
                    memcpy(recbuf.data() + at, &c, sizeof c);
                    memcpy(recbuf.data() + at + r.args_at, r.args.data(), r.args.size());
                }
                shm = std::max(shm, r.shm);
                nblk = next;
            }
            if (!nblk) continue;
            GroupLaunch g;
            g.kernel = r0.kernel;
            g.threads = r0.threads;
            g.shm = shm;
            g.nblocks = (uint32_t)nblk;
            g.recs_off = img.put(recbuf.data(), recbuf.size());
            if (shm > 64 * 1024)
                if (int s = bhip(b, hipFuncSetAttribute(g.kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm),
                                 "group LDS size"))
                    return s;
            out.push_back(g);
        }
    }
    return 0;
}

// Builds the group plan of a pair list (not under capture): sizes one
// workspace per chunk position of each slot, records every step of every
// eligible pair and uploads the tables.
int build_plan(apg_hip_batch b, const apg_hip_pair* pairs, int n) {
    drop_plan(b);
    const int nslots = (int)b->slots.size();
    const auto per = plan_slots(pairs, n, nslots);
    GroupPlan P;
    P.launches.resize(nslots);
    P.single.resize(nslots);
    if (b->gctx.size() < (size_t)nslots) b->gctx.resize(nslots);
    TableImage img;
    for (int s = 0; s < nslots; ++s) {
        hipStream_t st = (hipStream_t)apg_hip_stream(b->slots[s]);
        // chunks: the slot's pairs in order, up to kGroupChunkBytes of workspace each
        std::vector<std::vector<int>> chunks;
        size_t acc = 0;
        for (int i : per[s]) {
            const apg_hip_pair& q = pairs[i];
            const size_t states = (size_t)q.width * q.height * (size_t)(q.params.max_disparity - q.params.min_disparity);
            if (states > b->group_states) {
                P.single[s].push_back(i);
                continue;
            }
            const size_t ws = apg_hip_pipeline_workspace_bytes(&q.params, q.width, q.height);
            if (chunks.empty() || acc + ws > b->group_bytes) {
                chunks.emplace_back();
                acc = 0;
            }
            chunks.back().push_back(i);
            acc += ws;
        }
        size_t npos = 0;
        for (const auto& c : chunks) npos = std::max(npos, c.size());
        while (b->gctx[s].size() < npos) {
            apg_hip_ctx c = nullptr;
            if (int e = apgk_ctx_create_on(b->device, st, &c)) return bfail(b, e, "group context");
            b->gctx[s].push_back(c);
        }
        // every workspace grows to its size before any launch is recorded:
        // recorded launches hold the buffer addresses
        std::vector<std::vector<bool>> ok(chunks.size());
        for (size_t c = 0; c < chunks.size(); ++c)
            for (size_t k = 0; k < chunks[c].size(); ++k) {
                apg_hip_ctx ctx = b->gctx[s][k];
                const apg_hip_pair& q = pairs[chunks[c][k]];
                set_pair_tuning(b, ctx, q);
                const int e = apgk_group_prepare(ctx, q);
                if (e == 7) return bfail(b, 7, std::string("group workspace: ") + apg_hip_last_error(ctx));
                ok[c].push_back(e == 0);
            }
        for (size_t c = 0; c < chunks.size(); ++c) {
            // every step of each pair, recorded first, so that a pair with a
            // step that has no group form runs one at a time instead
            std::vector<std::vector<std::vector<apgk::GroupRec>>> rec(chunks[c].size());
            for (size_t k = 0; k < chunks[c].size(); ++k) {
                const int i = chunks[c][k];
                if (ok[c][k]) {
                    apg_hip_ctx ctx = b->gctx[s][k];
                    set_pair_tuning(b, ctx, pairs[i]);
                    rec[k].resize(kGroupSteps);
                    for (int step = 0; step < kGroupSteps && ok[c][k]; ++step) {
                        apgk::LaunchSink sink;
                        if (apgk_group_record(ctx, pairs[i], step, &sink)) ok[c][k] = false;
                        else rec[k][step] = std::move(sink.recs);
                    }
                }
                if (!ok[c][k]) P.single[s].push_back(i);
                else ++P.grouped;
            }
            for (int step = 0; step < kGroupSteps; ++step) {
                std::vector<apgk::GroupRec> recs;
                for (size_t k = 0; k < chunks[c].size(); ++k)
                    if (ok[c][k])
                        for (auto& r : rec[k][step]) recs.push_back(std::move(r));
                if (int e = group_step(b, recs, img, P.launches[s])) return e;
            }
        }
        P.nlaunch += (int)P.launches[s].size();
        P.nsingle += (int)P.single[s].size();
    }
    if (!img.bytes.empty()) {
        if (int e = bhip(b, hipMalloc(&P.d_tables, img.bytes.size()), "group tables")) return e;
        if (int e = bhip(b, hipMemcpy(P.d_tables, img.bytes.data(), img.bytes.size(), hipMemcpyHostToDevice),
                         "group tables upload")) {
            (void)hipFree(P.d_tables);
            return e;
        }
    }
    P.pairs.assign(pairs, pairs + n);
    P.valid = true;
    b->plan = std::move(P);
    return 0;
}

// Enqueues the plan: per slot its single pairs (the larger ones, first as in
// longest-processing-time order), then its group launches.
int enqueue_plan(apg_hip_batch b, const apg_hip_pair* pairs) {
    const GroupPlan& P = b->plan;
    if (int s = bhip(b, hipEventRecord(b->fork, b->origin), "fork record")) return s;
    int status = 0;
    for (size_t s = 0; s < b->slots.size(); ++s) {
        if (P.launches[s].empty() && P.single[s].empty()) continue;
        apg_hip_ctx ctx = b->slots[s];
        hipStream_t st = (hipStream_t)apg_hip_stream(ctx);
        if (int e = bhip(b, hipStreamWaitEvent(st, b->fork, 0), "fork wait")) return e;
        for (int i : P.single[s]) {
            if (status) break;
            set_pair_tuning(b, ctx, pairs[i]);
            status = run_pair(ctx, pairs[i]);
            if (status) b->last_error = "pair " + std::to_string(i) + ": " + apg_hip_last_error(ctx);
        }
        for (const GroupLaunch& g : P.launches[s]) {
            if (status) break;
            const void* rp = (const uint8_t*)P.d_tables + g.recs_off;
            void* kargs[] = {(void*)&rp};
            status = bhip(b, hipLaunchKernel(g.kernel, dim3(g.nblocks), dim3(g.threads), kargs, g.shm, st),
                          "group launch");
        }
        hipError_t e = hipEventRecord(b->join[s], st);
        if (e == hipSuccess) e = hipStreamWaitEvent(b->origin, b->join[s], 0);
        if (e != hipSuccess && !status) status = bhip(b, e, "join");
    }
    return status;
}

}  // namespace

extern "C" {

int apg_hip_batch_create(int device, int slots, apg_hip_batch* out) {
    if (!out) return 8;
    *out = nullptr;
    if (slots == 0) slots = 4;
    if (slots < 1 || slots > kMaxSlots) return 4;
    apg_hip_batch b = new (std::nothrow) apg_hip_batch_s();
    if (!b) return 7;
    b->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    int st = e == hipSuccess ? 0 : 8;
    for (int i = 0; i < slots && !st; ++i) {
        apg_hip_ctx c = nullptr;
        st = apg_hip_ctx_create(device, &c);
        if (!st) b->slots.push_back(c);
        hipEvent_t ev = nullptr;
        if (!st && hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) st = 8;
        if (!st) b->join.push_back(ev);
    }
    if (st) {
        apg_hip_batch_destroy(b);
        return st;
    }
    b->arch = device_arch(device);
    *out = b;
    return 0;
}

void apg_hip_batch_destroy(apg_hip_batch b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->origin) (void)hipStreamSynchronize(b->origin);
    drop_plan(b);
    for (auto& v : b->gctx)
        for (apg_hip_ctx c : v) apg_hip_ctx_destroy(c);
    for (apg_hip_ctx c : b->slots) apg_hip_ctx_destroy(c);
    for (hipEvent_t e : b->join) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->origin) (void)hipStreamDestroy(b->origin);
    delete b;
}

void* apg_hip_batch_stream(apg_hip_batch b) { return b ? (void*)b->origin : nullptr; }

int apg_hip_batch_synchronize(apg_hip_batch b) {
    if (!b) return 8;
    if (int s = bhip(b, hipStreamSynchronize(b->origin), "batch synchronize")) return s;
    // a HighQuality refiner inside a replayed graph cannot report a hand-off
    // wait that gave up when it runs: its slot's sticky status tells here
    for (size_t s = 0; s < b->slots.size(); ++s)
        if (int e = apgk_ctx_take_status(b->slots[s]))
            return bfail(b, e, "slot " + std::to_string(s) + ": " + apg_hip_last_error(b->slots[s]));
    return 0;
}

size_t apg_hip_batch_workspace_bytes(apg_hip_batch b) {
    if (!b) return 0;
    size_t t = 0;
    for (apg_hip_ctx c : b->slots) t += apg_hip_workspace_bytes(c);
    for (const auto& v : b->gctx)
        for (apg_hip_ctx c : v) t += apg_hip_workspace_bytes(c);
    return t;
}

const char* apg_hip_batch_last_error(apg_hip_batch b) { return b ? b->last_error.c_str() : "invalid batch"; }

int apg_hip_batch_run(apg_hip_batch b, const apg_hip_pair* pairs, int n, int flags) {
    if (!b) return 8;
    if (n < 0 || (n > 0 && !pairs)) return bfail(b, 4, "invalid pair list");
    if (n == 0) return 0;
    if (int s = bhip(b, hipSetDevice(b->device), "hipSetDevice")) return s;
    // (HighQuality pairs have no group form: they run one at a time on their
    // slot, in a captured graph like the others)
    const bool grouped = !(flags & APG_HIP_BATCH_PER_PAIR);
    if (!(flags & APG_HIP_BATCH_GRAPH)) {
        // a direct run may grow slot workspaces, which a cached graph points into
        drop_graph(b);
        if (!grouped) return enqueue(b, pairs, n);
        if (!b->plan.valid || !same_pairs(b->plan.pairs, pairs, n))
            if (int s = build_plan(b, pairs, n)) return s;
        return enqueue_plan(b, pairs);
    }
    if (b->exec && same_pairs(b->graph_pairs, pairs, n) && b->graph_grouped == grouped)
        return bhip(b, hipGraphLaunch(b->exec, b->origin), "hipGraphLaunch");
    drop_graph(b);
    if (grouped && (!b->plan.valid || !same_pairs(b->plan.pairs, pairs, n)))
        if (int s = build_plan(b, pairs, n)) return s;
    if (int s = grouped ? enqueue_plan(b, pairs) : enqueue(b, pairs, n)) return s;
    // The direct run sized every workspace and cached every LUT, so the
    // capture below records kernels only.  It is not executed: the next call
    // with this pair list replays it.
    if (int s = bhip(b, hipStreamBeginCapture(b->origin, hipStreamCaptureModeRelaxed), "begin capture")) return s;
    for (apg_hip_ctx c : b->slots) apgk_ctx_set_capturing(c, true);
    const int st = grouped ? enqueue_plan(b, pairs) : enqueue(b, pairs, n);
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(b->origin, &g);
    for (apg_hip_ctx c : b->slots) apgk_ctx_set_capturing(c, false);
    if (st || ee != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return st ? st : bhip(b, ee, "end capture");
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
        (void)hipGraphDestroy(g);
        return bhip(b, ie, "hipGraphInstantiate");
    }
    b->graph = g;
    b->exec = ex;
    b->graph_pairs.assign(pairs, pairs + n);
    b->graph_grouped = grouped;
    return 0;
}

int apg_hip_batch_set_grouping(apg_hip_batch b, size_t chunk_bytes, size_t max_pair_states) {
    if (!b) return 8;
    b->group_bytes = chunk_bytes ? chunk_bytes : kGroupChunkBytes;
    b->group_states = max_pair_states ? max_pair_states : kGroupMaxStates;
    drop_plan(b);
    return 0;
}

int apg_hip_batch_group_stats(apg_hip_batch b, int* grouped_pairs, int* group_launches, int* single_pairs) {
    if (!b) return 8;
    if (grouped_pairs) *grouped_pairs = b->plan.valid ? b->plan.grouped : 0;
    if (group_launches) *group_launches = b->plan.valid ? b->plan.nlaunch : 0;
    if (single_pairs) *single_pairs = b->plan.valid ? b->plan.nsingle : 0;
    return 0;
}

int apg_hip_batch_tuning_for(apg_hip_batch b, const apg_hip_params* p, int w, int h, apg_hip_tuning* out) {
    if (!b || !p || !out) return 8;
    const auto it = b->tuned.find(bucket_of(*p, w, h));
    *out = it == b->tuned.end() ? apg_hip_tuning{0, 0, 0} : it->second;
    return 0;
}

}  // extern "C"

// ============================================================= autotuner ====

namespace {

// Per-stage minima over `reps` timed runs (after one untimed run) of pair q
// on ctx with launch shapes t.
int time_shapes(apg_hip_ctx ctx, const apg_hip_pair& q, const apg_hip_tuning& t, int reps, apg_hip_timing* best) {
    if (int s = apg_hip_set_tuning(ctx, &t)) return s;
    apg_hip_enable_stage_timing(ctx, 0);
    if (int s = run_pair(ctx, q)) return s;
    apg_hip_enable_stage_timing(ctx, 1);
    apg_hip_timing m{};
    for (int r = 0; r < reps; ++r) {
        if (int s = run_pair(ctx, q)) return s;
        apg_hip_timing x{};
        apg_hip_last_timing(ctx, &x);
        if (r == 0 || x.cost_ms < m.cost_ms) m.cost_ms = x.cost_ms;
        if (r == 0 || x.cross_ms < m.cross_ms) m.cross_ms = x.cross_ms;
        if (r == 0 || x.sgm_ms < m.sgm_ms) m.sgm_ms = x.sgm_ms;
        if (r == 0 || x.wta_ms < m.wta_ms) m.wta_ms = x.wta_ms;
        if (r == 0 || x.total_ms < m.total_ms) m.total_ms = x.total_ms;
    }
    apg_hip_enable_stage_timing(ctx, 0);
    *best = m;
    return 0;
}

// Coordinate search: each knob is chosen from the time of the stages it
// changes, with the other knobs at their defaults: the cost segment from the
// cost pass, the Cross-V wave count from Cross-V, the SGM wave shape from the
// SGM plus the WTA (shapes that cannot pair a scanline family's planes leave
// the WTA more planes to read).  A candidate must beat the built-in shape by
// 2% to replace it.
int tune_pair(apg_hip_ctx ctx, const apg_hip_pair& q, apg_hip_tuning* out) {
    const int reps = 3;
    apg_hip_timing base{}, t{};
    if (int s = time_shapes(ctx, q, apg_hip_tuning{0, 0, 0}, reps, &base)) return s;
    apg_hip_tuning r{0, 0, 0};
    double bc = base.cost_ms, bv = base.cross_ms, bs = base.sgm_ms;
    for (int seg : {64, 128, 512}) {
        if (int s = time_shapes(ctx, q, apg_hip_tuning{seg, 0, 0}, reps, &t)) return s;
        if (t.cost_ms < 0.98 * bc) bc = t.cost_ms, r.cost_segment = seg;
    }
    for (int cv : {1024, 2048, 4096, 16384}) {
        if (int s = time_shapes(ctx, q, apg_hip_tuning{0, 0, cv}, reps, &t)) return s;
        if (t.cross_ms < 0.98 * bv) bv = t.cross_ms, r.cross_v_waves = cv;
    }
    bs += base.wta_ms;
    for (int sd : {1, 2, 4}) {
        if (int s = time_shapes(ctx, q, apg_hip_tuning{0, sd, 0}, reps, &t)) return s;
        if (t.sgm_ms + t.wta_ms < 0.98 * bs) bs = t.sgm_ms + t.wta_ms, r.sgm_dual = sd;
    }
    *out = r;
    return 0;
}

}  // namespace

extern "C" int apg_hip_batch_autotune(apg_hip_batch b, const apg_hip_pair* pairs, int n, const char* cache_path) {
    if (!b) return 8;
    if (n < 0 || (n > 0 && !pairs)) return bfail(b, 4, "invalid pair list");
    if (int s = bhip(b, hipSetDevice(b->device), "hipSetDevice")) return s;
    if (cache_path) load_cache(cache_path, b->arch, b->tuned);
    drop_plan(b);
    // tuning runs on slot 0 after everything queued on the batch
    apg_hip_ctx ctx = b->slots[0];
    if (int s = bhip(b, hipStreamSynchronize(b->origin), "sync")) return s;
    int status = 0;
    for (int i = 0; i < n && !status; ++i) {
        const apg_hip_pair& q = pairs[i];
        const BucketKey k = bucket_of(q.params, q.width, q.height);
        if (b->tuned.count(k)) continue;
        apg_hip_tuning t{};
        status = tune_pair(ctx, q, &t);
        if (status) b->last_error = "autotune pair " + std::to_string(i) + ": " + apg_hip_last_error(ctx);
        else b->tuned[k] = t;
    }
    apg_hip_enable_stage_timing(ctx, 0);
    (void)apg_hip_set_tuning(ctx, nullptr);
    if (!status && cache_path && !save_cache(cache_path, b->arch, b->tuned))
        return bfail(b, 8, std::string("cannot write ") + cache_path);
    return status;
}

// ================================================================ server ====

namespace {

struct Slot {
    apg_hip_ctx ctx = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t done = nullptr;
    uint8_t* d_io = nullptr;   // left | right | disparity | confidence | reason | reliable
    uint8_t* h_io = nullptr;   // pinned staging of the same layout
    size_t io_bytes = 0;
    size_t ws_bytes = 0;       // workspace bytes the slot holds
    bool busy = false;
    uint64_t tag = 0;
    apg_hip_request req{};
};

struct IoLayout {
    size_t left, right, disp, conf, reason, rel, total;
};


IoLayout io_layout(const apg_hip_request& r) {
    const size_t n = (size_t)r.width * r.height;
    const size_t img = (size_t)r.stride * r.height;
    IoLayout L;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = (o + bytes + 255) & ~(size_t)255;
        return at;
    };
    L.left = take(img);
    L.right = take(img);
    L.disp = take(n * 4);
    L.conf = take(n * 4);
    L.reason = take(n);
    L.rel = take(n);
    L.total = o;
    return L;
}

}  // namespace

struct apg_hip_server_s {
    int device = 0;
    size_t budget = 0;
    std::vector<Slot> slots;
    std::deque<std::pair<apg_hip_request, uint64_t>> queue;
    std::string last_error;
    std::unique_ptr<apgk::CopyPool> copies;
};

namespace {

int sfail(apg_hip_server s, int code, const std::string& msg) {
    if (s) s->last_error = msg;
    return code;
}

size_t request_bytes(const apg_hip_request& r) {
    return apg_hip_pipeline_workspace_bytes(&r.params, r.width, r.height) + io_layout(r).total;
}

size_t resident(const apg_hip_server_s* s) {
    size_t t = 0;
    for (const Slot& sl : s->slots) t += sl.ws_bytes + sl.io_bytes;
    return t;
}

void release_slot(Slot& sl) {
    apg_hip_release_workspace(sl.ctx);
    if (sl.d_io) (void)hipFree(sl.d_io);
    if (sl.h_io) (void)hipHostFree(sl.h_io);
    sl.d_io = sl.h_io = nullptr;
    sl.io_bytes = sl.ws_bytes = 0;
}

// Starts request r on free slot sl: inputs -> pinned staging -> device,
// pipeline, results -> pinned staging; `done` fires when all of it is over.
int start(apg_hip_server s, Slot& sl, const apg_hip_request& r, uint64_t tag) {
    const IoLayout L = io_layout(r);
    if (sl.io_bytes < L.total) {
        if (sl.d_io) (void)hipFree(sl.d_io);
        if (sl.h_io) (void)hipHostFree(sl.h_io);
        sl.d_io = sl.h_io = nullptr;
        sl.io_bytes = 0;
        if (hipMalloc(&sl.d_io, L.total) != hipSuccess) return sfail(s, 7, "device i/o buffer");
        if (hipHostMalloc(&sl.h_io, L.total, hipHostMallocDefault) != hipSuccess) {
            (void)hipFree(sl.d_io);
            sl.d_io = nullptr;
            return sfail(s, 7, "pinned i/o buffer");
        }
        sl.io_bytes = L.total;
    }
    const size_t img = (size_t)r.stride * r.height;
    s->copies->copy({{sl.h_io + L.left, r.left, img}, {sl.h_io + L.right, r.right, img}});
    hipError_t e = hipMemcpyAsync(sl.d_io, sl.h_io, L.disp, hipMemcpyHostToDevice, sl.st);
    if (e != hipSuccess) return sfail(s, 8, std::string("upload: ") + hipGetErrorString(e));
    const int st = apg_hip_pipeline(sl.ctx, &r.params, sl.d_io + L.left, sl.d_io + L.right, r.width, r.height,
                                    r.stride, r.channels, (float*)(sl.d_io + L.disp), (float*)(sl.d_io + L.conf),
                                    sl.d_io + L.reason, sl.d_io + L.rel, nullptr);
    sl.ws_bytes = apg_hip_workspace_bytes(sl.ctx);
    if (st) return sfail(s, st, apg_hip_last_error(sl.ctx));
    e = hipMemcpyAsync(sl.h_io + L.disp, sl.d_io + L.disp, L.total - L.disp, hipMemcpyDeviceToHost, sl.st);
    if (e == hipSuccess) e = hipEventRecord(sl.done, sl.st);
    if (e != hipSuccess) return sfail(s, 8, std::string("download: ") + hipGetErrorString(e));
    sl.busy = true;
    sl.tag = tag;
    sl.req = r;
    return 0;
}

void finish(apg_hip_server s, Slot& sl) {
    const apg_hip_request& r = sl.req;
    const IoLayout L = io_layout(r);
    const size_t n = (size_t)r.width * r.height;
    std::vector<std::tuple<void*, const void*, size_t>> out{{r.disparity, sl.h_io + L.disp, n * 4}};
    if (r.confidence) out.emplace_back(r.confidence, sl.h_io + L.conf, n * 4);
    if (r.invalid_reason) out.emplace_back(r.invalid_reason, sl.h_io + L.reason, n);
    if (r.reliable) out.emplace_back(r.reliable, sl.h_io + L.rel, n);
    s->copies->copy(out);
    sl.busy = false;
}

// Device bytes slot sl would add to run request r.
size_t growth(const Slot& sl, const apg_hip_request& r) {
    const size_t io = io_layout(r).total;
    return apgk_ctx_growth_bytes(sl.ctx, &r.params, r.width, r.height) + (io > sl.io_bytes ? io - sl.io_bytes : 0);
}

// Picks a free slot for request r: the one that needs to grow least (the
// smallest among those that need not grow at all), provided the resident
// bytes stay within the budget, releasing idle workspaces (largest first)
// when that makes room.  nullptr: wait for a slot to finish.
Slot* admit_slot(apg_hip_server s, const apg_hip_request& r) {
    Slot* best = nullptr;
    size_t best_g = 0;
    for (Slot& sl : s->slots) {
        if (sl.busy) continue;
        const size_t g = growth(sl, r);
        if (!best || g < best_g || (g == best_g && sl.ws_bytes + sl.io_bytes < best->ws_bytes + best->io_bytes)) {
            best = &sl;
            best_g = g;
        }
    }
    if (!best) return nullptr;
    size_t held = resident(s);
    if (held + best_g <= s->budget) return best;
    std::vector<Slot*> idle;
    for (Slot& sl : s->slots)
        if (!sl.busy && &sl != best && sl.ws_bytes + sl.io_bytes > 0) idle.push_back(&sl);
    std::sort(idle.begin(), idle.end(),
              [](Slot* a, Slot* c) { return a->ws_bytes + a->io_bytes > c->ws_bytes + c->io_bytes; });
    for (Slot* sl : idle) {
        if (held + best_g <= s->budget) break;
        held -= sl->ws_bytes + sl->io_bytes;
        release_slot(*sl);
    }
    if (held + best_g > s->budget) {
        // the slot's own buffers may be shaped for other sizes: start it afresh
        held -= best->ws_bytes + best->io_bytes;
        release_slot(*best);
        best_g = growth(*best, r);
    }
    return held + best_g <= s->budget ? best : nullptr;
}

}  // namespace

extern "C" {

int apg_hip_server_create(int device, int slots, size_t hbm_budget_bytes, apg_hip_server* out) {
    if (!out) return 8;
    *out = nullptr;
    if (slots == 0) slots = 4;
    if (slots < 1 || slots > kMaxSlots) return 4;
    apg_hip_server s = new (std::nothrow) apg_hip_server_s();
    if (!s) return 7;
    s->device = device;
    s->budget = hbm_budget_bytes ? hbm_budget_bytes : ((size_t)64 << 30);
    s->slots.resize(slots);
    // staging copies: the polling thread and up to 7 more
    s->copies.reset(new (std::nothrow) apgk::CopyPool(apgk::CopyPool::default_threads()));
    int st = s->copies ? (hipSetDevice(device) == hipSuccess ? 0 : 8) : 7;
    for (Slot& sl : s->slots) {
        if (st) break;
        st = apg_hip_ctx_create(device, &sl.ctx);
        if (!st) sl.st = (hipStream_t)apg_hip_stream(sl.ctx);
        if (!st && hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) st = 8;
    }
    if (st) {
        apg_hip_server_destroy(s);
        return st;
    }
    *out = s;
    return 0;
}

void apg_hip_server_destroy(apg_hip_server s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    for (Slot& sl : s->slots) {
        if (sl.st) (void)hipStreamSynchronize(sl.st);
        if (sl.ctx) release_slot(sl);
        if (sl.done) (void)hipEventDestroy(sl.done);
        apg_hip_ctx_destroy(sl.ctx);
    }
    delete s;
}

const char* apg_hip_server_last_error(apg_hip_server s) { return s ? s->last_error.c_str() : "invalid server"; }

int apg_hip_server_submit(apg_hip_server s, const apg_hip_request* r, uint64_t tag) {
    if (!s) return 8;
    if (!r || !r->left || !r->right || !r->disparity || r->width <= 0 || r->height <= 0)
        return sfail(s, 1, "empty request");
    if ((r->channels != 1 && r->channels != 3) || r->stride < r->width * r->channels)
        return sfail(s, 2, "invalid row stride");
    if (int v = apg_hip_validate_params(&r->params)) return sfail(s, v, "invalid parameters");
    if (request_bytes(*r) > s->budget) return sfail(s, 6, "request exceeds the HBM budget");
    s->queue.emplace_back(*r, tag);
    return 0;
}

int apg_hip_server_poll(apg_hip_server s, uint64_t* done_tags, int* done_status, int max_done) {
    if (!s) return -1;
    if (hipSetDevice(s->device) != hipSuccess) return -1;
    int nd = 0;
    for (Slot& sl : s->slots) {
        if (!sl.busy || nd >= max_done) continue;
        const hipError_t q = hipEventQuery(sl.done);
        if (q == hipErrorNotReady) continue;
        int st = 0;
        if (q == hipSuccess) finish(s, sl);
        else st = sfail(s, 8, std::string("request: ") + hipGetErrorString(q));
        sl.busy = false;
        done_tags[nd] = sl.tag;
        if (done_status) done_status[nd] = st;
        ++nd;
    }
    // admission in arrival order; a request that fails to start is reported
    // as done with its status
    while (!s->queue.empty() && nd < max_done) {
        const apg_hip_request& r = s->queue.front().first;
        Slot* sl = admit_slot(s, r);
        if (!sl) break;
        const uint64_t tag = s->queue.front().second;
        const int st = start(s, *sl, r, tag);
        if (st) {
            done_tags[nd] = tag;
            if (done_status) done_status[nd] = st;
            ++nd;
        }
        s->queue.pop_front();
    }
    return nd;
}

int apg_hip_server_pending(apg_hip_server s) {
    if (!s) return 0;
    int n = (int)s->queue.size();
    for (const Slot& sl : s->slots) n += sl.busy ? 1 : 0;
    return n;
}

size_t apg_hip_server_resident_bytes(apg_hip_server s) { return s ? resident(s) : 0; }

int apg_hip_server_replay(apg_hip_server s, const apg_hip_request* reqs, const double* arrival_s, int n,
                          double* latency_s) {
    if (!s) return 8;
    if (n < 0 || (n > 0 && (!reqs || !arrival_s || !latency_s))) return sfail(s, 4, "invalid trace");
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    auto now = [&] { return std::chrono::duration<double>(clk::now() - t0).count(); };
    std::vector<uint64_t> tags(kMaxSlots + 64);
    std::vector<int> sts(kMaxSlots + 64);
    int next = 0, done = 0, first_err = 0;
    while (done < n) {
        const double t = now();
        while (next < n && arrival_s[next] <= t) {
            const int st = apg_hip_server_submit(s, &reqs[next], (uint64_t)next);
            if (st) {
                latency_s[next] = -1.0;
                if (!first_err) first_err = st;
                ++done;
            }
            ++next;
        }
        const int k = apg_hip_server_poll(s, tags.data(), sts.data(), (int)tags.size());
        if (k < 0) return sfail(s, 8, "poll failed");
        const double tc = now();
        for (int i = 0; i < k; ++i) {
            const int id = (int)tags[i];
            latency_s[id] = sts[i] ? -1.0 : tc - arrival_s[id];
            if (sts[i] && !first_err) first_err = sts[i];
            ++done;
        }
        if (k == 0 && !s->queue.empty()) {
            bool busy = false;
            for (const Slot& sl : s->slots) busy = busy || sl.busy;
            if (!busy) return sfail(s, 8, "admission stalled with every slot idle");
        }
        if (k == 0 && next < n && arrival_s[next] - now() > 200e-6) std::this_thread::yield();
    }
    return first_err;
}

}  // extern "C"

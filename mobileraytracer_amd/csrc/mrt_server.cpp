// Render server: continuous batching of render jobs over resident scenes.
//
// The reference renders one frame per Renderer::renderFrame call
// (Renderer.cpp:31-59); a server runs many requests at once through the
// wavefront pipeline of mrt_wavefront.hip:
//   * scenes are made resident once: their walk trees (pair tree, compressed
//     wide tree, leaf records and triangle slots) are copied into one arena
//     with node and slot references rebased, so one persistent trace launch
//     walks rays of every scene, each from its scene's roots;
//   * the path-slot pool is cut into pipelines (own queues and stream), and
//     the pool's 64-slot waves are handed to jobs as they are admitted; each
//     wave records its job, so the regen and shade kernels of a pipeline serve
//     every job on it, each slot with its own job's parameters and scene;
//   * a worker thread issues the pipelines' iterations in batches, admits
//     queued jobs between batches (in submission order, as waves, a job-table
//     entry and job memory are free), and reads per-job finished-slot counters
//     (one array per pipeline, copied on the pipeline's own stream) one batch
//     behind; a job whose every covered pixel finished its last sample is
//     copied out and its waves go to the next job.
// The pipelines are the context's slot-group streams and never wait on one
// another: a stream that waits on the others' events stalls every stream
// that shares its hardware queue.  Job inputs go up by DMA from pinned
// staging before the admit kernels are queued, and results come back once
// the counters show the job done, which means the kernels that wrote them have
// completed.
// A job's pixels take the same draws and the same arithmetic as in mrt_render,
// so its bitmap and accumulation are identical to rendering it alone.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <unordered_map>

#include "mrt_internal.hpp"

using namespace mrt;

namespace {

constexpr uint32_t kBatch = 4;  // iterations per pipeline between counter snapshots
constexpr uint32_t kDefaultSlots = 33554432u;
constexpr uint32_t kDefaultJobs = 256u;
constexpr uint32_t kDefaultScenes = 64u;
constexpr uint64_t kDefaultJobMemory = 1ull << 30;
constexpr uint64_t kDefaultSampleMemory = 8ull << 30;
constexpr uint32_t kMaxJobLanes = 8;
constexpr uint64_t kAlign = 256;

double now_us() {
    return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

uint64_t align_up(uint64_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

// First-fit allocator over [0, size): free ranges keyed by their start.
class RangeAlloc {
  public:
    void reset(uint64_t begin, uint64_t size) {
        free_.clear();
        total_ = size;
        if (size) free_[begin] = size;
    }
    uint64_t free_total() const {
        uint64_t t = 0;
        for (const auto& kv : free_) t += kv.second;
        return t;
    }
    uint64_t total() const { return total_; }
    // One contiguous range of n units.
    bool alloc(uint64_t n, uint64_t* off) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            *off = it->first;
            const uint64_t rest = it->second - n, b = it->first + n;
            free_.erase(it);
            if (rest) free_[b] = rest;
            return true;
        }
        return false;
    }
    // n units from the lowest ranges, as (start, count) runs.
    void take(uint64_t n, std::vector<std::pair<uint64_t, uint64_t>>* runs) {
        while (n > 0 && !free_.empty()) {
            auto it = free_.begin();
            const uint64_t k = std::min(n, it->second);
            runs->emplace_back(it->first, k);
            const uint64_t rest = it->second - k, b = it->first + k;
            free_.erase(it);
            if (rest) free_[b] = rest;
            n -= k;
        }
    }
    void release(uint64_t off, uint64_t n) {
        if (n == 0) return;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + n == next->first) {
            n += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += n;
                return;
            }
        }
        free_[off] = n;
    }

  private:
    std::map<uint64_t, uint64_t> free_;
    uint64_t total_ = 0;
};

struct Run {
    uint32_t pipe, pool_wave, job_wave, n;
};

enum JobPhase { kQueued, kRunning, kCopying, kDone, kFailed };

struct Job {
    uint64_t id = 0;
    uint32_t scene = 0;
    mrt_render_desc desc{};
    RenderParams p{};
    std::vector<uint32_t> rank;
    uint32_t* bitmap = nullptr;  // caller's
    float* accum = nullptr;
    int mem = MRT_MEM_HOST;
    size_t bm_bytes = 0, acc_bytes = 0;
    uint32_t waves = 0, live = 0;     // admitted: waves_px * lanes waves, covered pixels * lanes slots
    uint32_t waves_px = 0;            // waves of one slot per pixel
    uint32_t lanes = 1;               // slots per pixel (sample-parallel jobs: > 1)
    uint64_t rad_off = 0, rad_units = 0;  // sample-memory block: done word, then per-sample radiance
    JobPhase phase = kQueued;
    int status = MRT_OK;
    std::string error;
    uint32_t slot = 0;                // job-table entry
    uint64_t first_snap = 0;          // first counter snapshot taken after admission
    uint64_t mem_bytes = 0;           // job memory block: bytes, and in allocator units
    uint64_t mem_off = 0, mem_units = 0;
    uint64_t off_bm = 0, off_acc = 0;  // inside the job memory block
    std::vector<Run> runs;
    uint32_t pipes = 0;               // bit g: the job has waves on pipeline g
    uint32_t stride = 0;              // caller's bitmap row stride (bytes)
    bool direct = false;              // rendered alone through mrt_render by an idle server
    hipEvent_t done_ev = nullptr;
    bool cancel = false;
    double t_submit = 0, t_admit = 0, t_gpu = 0, t_done = 0;
};

struct Resident {
    mrt_scene* scene;
    uint32_t wide_base, walk_base, tri_base;
};

}  // namespace

struct mrt_server {
    mrt_ctx* ctx = nullptr;
    uint32_t G = 0, max_jobs = 0, max_scenes = 0, pool_waves = 0;
    uint64_t job_memory = 0, scene_memory = 0;
    hipStream_t pipe[kWfMaxGroups] = {};  // the context's slot-group streams
    hipStream_t xfer = nullptr;           // result copies
    hipEvent_t chk[2][kWfMaxGroups] = {};
    // Device state.
    DevBuf state, group[kWfMaxGroups], jobs_dev, scenes_dev, waves_dev, finished_dev, counters, err, spill, job_mem;
    DevBuf arena_cw, arena_recs, arena_walk, arena_geo;
    uint64_t wide_used = 0, walk_used = 0, geo_used = 0;  // in nodes / pairs / slots
    uint32_t depth_need = 0;
    WfBuffers wb[kWfMaxGroups]{};
    WfLoop L[kWfMaxGroups]{};
    StackCfg sc[kWfMaxGroups]{};
    uint32_t trace_blocks = 0, persist_blocks = 0;
    // Pinned host memory: job table mirror, counter snapshots, job staging.
    JobDev* jobs_host = nullptr;
    uint32_t* snap_host = nullptr;  // [2][G][max_jobs + 1]: finished counters, then the error word
    uint64_t snap_batch[2] = {};    // batch number of each snapshot
    uint32_t snap_mask[2] = {};     // pipelines each snapshot covers
    uint8_t* staging = nullptr;
    // Host bookkeeping (guarded by mu).
    std::mutex mu;
    std::condition_variable cv, done_cv;
    std::vector<Resident> scenes;
    RangeAlloc waves_free[kWfMaxGroups];
    RangeAlloc mem_free;
    // Per-sample radiance of sample-parallel jobs (device only).
    DevBuf sample_mem;
    RangeAlloc sample_free;
    std::vector<uint32_t> free_slots;
    std::unordered_map<uint64_t, std::unique_ptr<Job>> jobs;
    std::deque<uint64_t> queue;
    std::vector<uint64_t> running;  // jobs in kRunning or kCopying
    uint64_t next_id = 1;
    uint64_t submitted = 0, done = 0, failed = 0, batches = 0;
    bool shutdown = false;
    bool broken = false;
    // Idle-direct renders: on, the job being rendered (0: none), and the rays
    // of those renders (camera, extension, shadow).
    bool idle_direct = false;
    uint64_t direct_job = 0;
    uint64_t direct_rays[3] = {};
    std::string broken_msg;
    std::thread worker;
};

namespace {

bool pipe_active(const mrt_server* s, uint32_t g) {
    return s->waves_free[g].free_total() < s->waves_free[g].total();
}

BatchDev batch_dev(const mrt_server* s, uint32_t g) {
    BatchDev bd;
    bd.jobs = s->jobs_dev.as<JobDev>();
    bd.scenes = s->scenes_dev.as<SceneDev>();
    bd.waves = s->waves_dev.as<uint2>();
    bd.finished = s->finished_dev.as<uint32_t>() + static_cast<size_t>(g) * s->max_jobs;
    return bd;
}

uint32_t* snap_of(const mrt_server* s, uint32_t snap, uint32_t g) {
    return s->snap_host + (static_cast<size_t>(snap) * s->G + g) * (s->max_jobs + 1u);
}

// The arena as a scene for the trace kernel: walk trees and triangle slots.
SceneDev arena_scene(const mrt_server* s) {
    SceneDev a{};
    a.walk_nodes = s->arena_walk.as<BNode>();
    a.cw_nodes = s->arena_cw.as<uint4>();
    a.cw_recs = s->arena_recs.as<float4>();
    a.tri_geo = s->arena_geo.as<float4>();
    return a;
}

// Waits for every stream of the server (arena or spill reallocation).
int quiesce(mrt_server* s) {
    for (uint32_t g = 0; g < s->G; ++g) MRT_HIP(hipStreamSynchronize(s->pipe[g]));
    MRT_HIP(hipStreamSynchronize(s->xfer));
    return MRT_OK;
}

// Grows a device buffer to `bytes`, keeping its first `keep` bytes.
int grow(mrt_server* s, DevBuf& b, size_t bytes, size_t keep) {
    if (bytes <= b.bytes) return MRT_OK;
    int rc;
    if ((rc = quiesce(s)) != MRT_OK) return rc;
    DevBuf nb;
    if ((rc = nb.alloc(bytes)) != MRT_OK) return rc;
    if (keep) MRT_HIP(hipMemcpy(nb.p, b.p, keep, hipMemcpyDeviceToDevice));
    std::swap(b.p, nb.p);
    std::swap(b.bytes, nb.bytes);
    return MRT_OK;
}

// Spill areas of the pipelines' traversal stacks for the deepest resident tree.
int ensure_spill(mrt_server* s, uint32_t depth_need) {
    const uint32_t need = depth_need + 1u;
    const uint32_t cap = need > kStackLds ? need - kStackLds : 0u;
    const uint32_t lanes = s->persist_blocks * 256u;
    if (cap <= s->sc[0].cap_spill && (cap == 0 || s->spill.p)) return MRT_OK;
    int rc;
    if ((rc = quiesce(s)) != MRT_OK) return rc;
    if ((rc = s->spill.alloc(static_cast<size_t>(cap) * lanes * s->G * sizeof(uint32_t))) != MRT_OK) return rc;
    for (uint32_t g = 0; g < s->G; ++g) {
        s->sc[g].cap_spill = cap;
        s->sc[g].spill_stride = lanes;
        s->sc[g].spill = s->spill.as<uint32_t>() + static_cast<size_t>(g) * cap * lanes;
    }
    return MRT_OK;
}

void fail_job(mrt_server* s, Job* j, int status, const std::string& msg) {
    j->phase = kFailed;
    j->status = status;
    j->error = msg;
    j->t_done = now_us();
    ++s->failed;
    s->done_cv.notify_all();
}

// Frees what an ended job holds on the device and in the tables.
void release_job(mrt_server* s, Job* j) {
    if (j->mem_units) s->mem_free.release(j->mem_off, j->mem_units);
    j->mem_units = 0;
    if (j->rad_units) s->sample_free.release(j->rad_off, j->rad_units);
    j->rad_units = 0;
    s->free_slots.push_back(j->slot);
    if (j->done_ev) (void)hipEventDestroy(j->done_ev);
    j->done_ev = nullptr;
}

// Admits queued jobs in submission order while waves, a table entry and job
// memory are free.  Called by the worker with mu held.
int admit(mrt_server* s) {
    std::vector<Job*> admitted;
    while (!s->queue.empty()) {
        Job* j = s->jobs[s->queue.front()].get();
        uint64_t freeWaves = 0;
        for (uint32_t g = 0; g < s->G; ++g) freeWaves += s->waves_free[g].free_total();
        if (freeWaves < j->waves_px || s->free_slots.empty()) break;
        uint64_t off = 0;
        if (!s->mem_free.alloc(j->mem_bytes / kAlign, &off)) break;
        j->mem_off = off;
        j->mem_units = j->mem_bytes / kAlign;
        // Sample-parallel job: up to 8 slots per pixel while they take at
        // most half of the free waves and the per-sample radiance fits.
        const uint32_t samples = j->desc.samples_pixel - j->desc.first_sample;
        uint32_t lanes = std::min(kMaxJobLanes, samples);
        while (lanes > 1u && static_cast<uint64_t>(j->waves_px) * lanes * 2u > freeWaves) --lanes;
        if (lanes > 1u) {
            const uint64_t units = align_up(kAlign + 12ull * samples * j->waves_px * 64u) / kAlign;
            uint64_t roff = 0;
            if (s->sample_free.alloc(units, &roff)) {
                j->rad_off = roff;
                j->rad_units = units;
            } else {
                lanes = 1u;
            }
        }
        j->lanes = lanes;
        j->waves = j->waves_px * lanes;
        j->live = 256u * j->p.bsx * j->p.bsy * lanes;
        s->queue.pop_front();
        j->slot = s->free_slots.back();
        s->free_slots.pop_back();
        // Waves: each pipeline's share in proportion to its free waves.
        std::vector<uint64_t> share(s->G, 0);
        uint64_t left = j->waves;
        for (uint32_t g = 0; g < s->G; ++g) {
            const uint64_t f = s->waves_free[g].free_total();
            share[g] = std::min<uint64_t>(f, (j->waves * f + freeWaves - 1) / freeWaves);
        }
        for (uint32_t g = 0; g < s->G; ++g) {
            share[g] = std::min(share[g], left);
            left -= share[g];
        }
        for (uint32_t g = 0; g < s->G && left > 0; ++g) {
            const uint64_t extra = std::min(left, s->waves_free[g].free_total() - share[g]);
            share[g] += extra;
            left -= extra;
        }
        uint32_t jw = 0;
        for (uint32_t g = 0; g < s->G; ++g) {
            std::vector<std::pair<uint64_t, uint64_t>> runs;
            s->waves_free[g].take(share[g], &runs);
            for (const auto& r : runs) {
                j->runs.push_back(Run{g, static_cast<uint32_t>(r.first), jw, static_cast<uint32_t>(r.second)});
                jw += static_cast<uint32_t>(r.second);
            }
        }
        j->phase = kRunning;
        j->t_admit = now_us();
        // Snapshots taken before this admission may hold the table entry's
        // previous count.
        j->first_snap = s->batches;
        s->running.push_back(j->id);
        admitted.push_back(j);
    }
    if (admitted.empty()) return MRT_OK;
    uint8_t* dmem = s->job_mem.as<uint8_t>();
    for (Job* j : admitted) {
        const uint64_t base = j->mem_off * kAlign;
        uint8_t* host = s->staging + base;
        std::memcpy(host, j->rank.data(), j->rank.size() * sizeof(uint32_t));
        JobDev jd{};
        jd.p = j->p;
        jd.p.tile_rank = reinterpret_cast<const uint32_t*>(dmem + base);
        jd.first_sample = j->desc.first_sample;
        jd.spp = j->desc.samples_pixel;
        jd.stop = 0u;
        jd.scene = j->scene;
        jd.wide_root = s->scenes[j->scene].wide_base;
        jd.walk_root = s->scenes[j->scene].walk_base;
        jd.rad = nullptr;
        jd.lanes = j->lanes;
        jd.pixel_slots = j->waves_px * 64u;
        if (j->lanes > 1u) {
            uint8_t* block = s->sample_mem.as<uint8_t>() + j->rad_off * kAlign;
            jd.rad = reinterpret_cast<float*>(block + kAlign);
            const size_t words = 3ull * (j->desc.samples_pixel - j->desc.first_sample) * jd.pixel_slots;
            MRT_HIP(hipMemsetD32(reinterpret_cast<hipDeviceptr_t>(jd.rad), static_cast<int>(kRadPending), words));
        }
        if (j->mem == MRT_MEM_DEVICE) {
            jd.bitmap = j->bitmap;
            jd.accum = j->accum;
        } else {
            std::memcpy(host + j->off_bm, j->bitmap, j->bm_bytes);
            jd.bitmap = reinterpret_cast<uint32_t*>(dmem + base + j->off_bm);
            jd.accum = nullptr;
            if (j->accum) {
                std::memcpy(host + j->off_acc, j->accum, j->acc_bytes);
                jd.accum = reinterpret_cast<float*>(dmem + base + j->off_acc);
            }
        }
        s->jobs_host[j->slot] = jd;
        // DMA from pinned staging, ordered before the admit kernels by the
        // host; no pipeline stream waits for it.
        const size_t upBytes = j->mem == MRT_MEM_DEVICE ? j->rank.size() * sizeof(uint32_t) : j->mem_bytes;
        MRT_HIP(hipMemcpy(dmem + base, host, upBytes, hipMemcpyHostToDevice));
        for (const Run& r : j->runs) {
            const bool first = !(j->pipes & (1u << r.pipe));
            j->pipes |= 1u << r.pipe;
            MRT_HIP(launch_wavefront_admit(batch_dev(s, r.pipe), jd, j->slot, r.pool_wave, r.job_wave, r.n, first,
                                           s->wb[r.pipe], s->L[r.pipe], s->pipe[r.pipe]));
        }
    }
    return MRT_OK;
}

// Issues one batch of iterations on every pipeline holding jobs and a
// snapshot of the finished counters.  Returns whether anything was issued.
int issue_batch(mrt_server* s, uint32_t snap, bool* issued) {
    *issued = false;
    const SceneDev arena = arena_scene(s);
    RenderParams none{};
    uint32_t mask = 0;
    for (uint32_t g = 0; g < s->G; ++g)
        if (pipe_active(s, g)) mask |= 1u << g;
    if (!mask) return MRT_OK;
    // Launches sized by the pipeline's slots that hold jobs (admitted before
    // this batch; finished ones not yet released only make it larger).
    uint32_t tb[kWfMaxGroups] = {}, pb[kWfMaxGroups] = {};
    for (uint32_t g = 0; g < s->G; ++g) {
        const uint64_t used = 64u * (s->waves_free[g].total() - s->waves_free[g].free_total());
        tb[g] = sized_blocks(s->trace_blocks, used);
        pb[g] = sized_trace_blocks(s->persist_blocks, used);
        s->wb[g].overlap = wavefront_overlap(2u * used, s->persist_blocks, s->ctx->tuning.overlap_loads) ? 1u : 0u;
    }
    for (uint32_t k = 0; k < kBatch; ++k) {
        for (uint32_t g = 0; g < s->G; ++g) {
            if (!(mask & (1u << g))) continue;
            const BatchDev bd = batch_dev(s, g);
            MRT_HIP(launch_wavefront_iteration(arena, false, true, false, none, 0u, false, nullptr, nullptr, s->wb[g],
                                               &s->L[g], tb[g], pb[g], s->sc[g],
                                               s->counters.as<unsigned long long>(), s->err.as<uint32_t>(), nullptr,
                                               nullptr, s->pipe[g], &bd));
        }
    }
    for (uint32_t g = 0; g < s->G; ++g) {
        if (!(mask & (1u << g))) continue;
        uint32_t* dst = snap_of(s, snap, g);
        MRT_HIP(hipMemcpyAsync(dst, s->finished_dev.as<uint32_t>() + static_cast<size_t>(g) * s->max_jobs,
                               s->max_jobs * sizeof(uint32_t), hipMemcpyDeviceToHost, s->pipe[g]));
        MRT_HIP(hipMemcpyAsync(dst + s->max_jobs, s->err.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s->pipe[g]));
        MRT_HIP(hipEventRecord(s->chk[snap][g], s->pipe[g]));
    }
    s->snap_batch[snap] = s->batches;
    s->snap_mask[snap] = mask;
    ++s->batches;
    *issued = true;
    return MRT_OK;
}

// Jobs whose every covered pixel finished in snapshot `snap`: free their
// waves and queue the copy of their outputs.
int process_snapshot(mrt_server* s, uint32_t snap) {
    const uint32_t mask = s->snap_mask[snap];
    for (uint32_t g = 0; g < s->G; ++g) {
        if ((mask & (1u << g)) && snap_of(s, snap, g)[s->max_jobs] != 0u && !s->broken) {
            s->broken = true;
            s->broken_msg = "traversal stack overflow in the render server";
        }
    }
    uint8_t* dmem = s->job_mem.as<uint8_t>();
    for (uint64_t id : s->running) {
        Job* j = s->jobs[id].get();
        if (j->phase != kRunning || s->snap_batch[snap] < j->first_snap || (j->pipes & ~mask) != 0u) continue;
        uint64_t fin = 0;
        for (uint32_t g = 0; g < s->G; ++g)
            if (j->pipes & (1u << g)) fin += snap_of(s, snap, g)[j->slot];
        if (fin < j->live) continue;
        j->phase = kCopying;
        j->t_gpu = now_us();
        for (const Run& r : j->runs) s->waves_free[r.pipe].release(r.pool_wave, r.n);
        j->runs.clear();
        if (j->lanes > 1u) {
            // Sample-parallel job: fold its samples into the outputs in
            // sample order (the kernels that wrote them have completed).
            const JobDev& jd = s->jobs_host[j->slot];
            WfBuffers wb{};
            wb.rad = jd.rad;
            wb.lanes = jd.lanes;
            wb.pixel_slots = jd.pixel_slots;
            wb.first = jd.first_sample;
            uint32_t* doneMin = reinterpret_cast<uint32_t*>(s->sample_mem.as<uint8_t>() + j->rad_off * kAlign);
            MRT_HIP(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(doneMin), static_cast<int>(jd.spp), 1, s->xfer));
            MRT_HIP(launch_wavefront_resolve(jd.p, jd.spp, wb, jd.bitmap, jd.accum, doneMin, s->xfer));
        }
        // The snapshot was copied after the kernels that finished the job,
        // and those have completed: the copy needs no stream dependency.
        if (j->mem != MRT_MEM_DEVICE) {
            const uint64_t base = j->mem_off * kAlign;
            MRT_HIP(hipMemcpyAsync(s->staging + base + j->off_bm, dmem + base + j->off_bm, j->bm_bytes,
                                   hipMemcpyDeviceToHost, s->xfer));
            if (j->accum)
                MRT_HIP(hipMemcpyAsync(s->staging + base + j->off_acc, dmem + base + j->off_acc, j->acc_bytes,
                                       hipMemcpyDeviceToHost, s->xfer));
        }
        MRT_HIP(hipEventCreateWithFlags(&j->done_ev, hipEventDisableTiming));
        MRT_HIP(hipEventRecord(j->done_ev, s->xfer));
    }
    return MRT_OK;
}

// Jobs whose output copy has landed: hand the result over.
int retire(mrt_server* s) {
    std::vector<uint64_t> keep;
    for (uint64_t id : s->running) {
        Job* j = s->jobs[id].get();
        if (j->phase == kCopying) {
            const hipError_t q = hipEventQuery(j->done_ev);
            if (q == hipErrorNotReady) {
                keep.push_back(id);
                continue;
            }
            if (q != hipSuccess) return fail(MRT_E_HIP, std::string("hipEventQuery: ") + hipGetErrorString(q));
            if (j->mem != MRT_MEM_DEVICE) {
                const uint64_t base = j->mem_off * kAlign;
                std::memcpy(j->bitmap, s->staging + base + j->off_bm, j->bm_bytes);
                if (j->accum) std::memcpy(j->accum, s->staging + base + j->off_acc, j->acc_bytes);
            }
            release_job(s, j);
            if (s->broken) {
                fail_job(s, j, MRT_E_INVALID, s->broken_msg);
                continue;
            }
            if (j->cancel) {
                // Stopped: the samples done so far are in the caller's buffers.
                fail_job(s, j, MRT_E_STOPPED, "render stopped");
                continue;
            }
            j->phase = kDone;
            j->status = MRT_OK;
            j->t_done = now_us();
            ++s->done;
            s->done_cv.notify_all();
            continue;
        }
        keep.push_back(id);
    }
    s->running.swap(keep);
    return MRT_OK;
}

// A job an idle server renders alone: the only job in the server (none
// admitted, none queued behind it), and one that mrt_render's AUTO gives the
// one-launch persistent megakernel (below kMegaMaxPixelSamples pixel samples),
// which finishes a small frame sooner than the pipelines' iterations.  Jobs
// that queue up together go to the pipelines.
bool direct_fits(const mrt_server* s, const Job* j) {
    if (!s->idle_direct || j->desc.persistent != 0u) return false;
    const uint64_t px = 256ull * j->p.bsx * j->p.bsy;  // every tile (servers refuse tile_skip)
    return px * (j->desc.samples_pixel - j->desc.first_sample) < kMegaMaxPixelSamples;
}

// Renders the queue's head alone through mrt_render, without the server's
// lock (submissions, polls and cancels go on meanwhile; the pipelines are
// idle).  Returns a device error that ends the server, else MRT_OK (a job
// that failed otherwise has failed on its own).
int run_direct(mrt_server* s, Job* j, std::unique_lock<std::mutex>& lk) {
    s->queue.pop_front();
    j->phase = kRunning;
    j->direct = true;
    j->lanes = 1u;
    j->t_admit = now_us();
    s->direct_job = j->id;
    mrt_scene* scene = s->scenes[j->scene].scene;
    mrt_render_desc d = j->desc;
    d.kernel = MRT_KERNEL_AUTO;
    mrt_render_stats st{};
    lk.unlock();
    const int rc = mrt_render(scene, &d, j->bitmap, j->stride, j->accum, j->mem, &st);
    const std::string err = rc == MRT_OK ? std::string() : g_error;
    lk.lock();
    s->direct_job = 0;
    j->t_gpu = now_us();
    if (rc != MRT_OK) {
        fail_job(s, j, rc, err);
        if (rc == MRT_E_HIP) {
            g_error = err;
            return rc;
        }
        return MRT_OK;
    }
    s->direct_rays[0] += st.rays_camera;
    s->direct_rays[1] += st.rays_extension;
    s->direct_rays[2] += st.rays_shadow;
    if (j->cancel) {
        // Stopped while rendering: every sample is in the caller's buffers.
        fail_job(s, j, MRT_E_STOPPED, "render stopped");
        return MRT_OK;
    }
    j->phase = kDone;
    j->status = MRT_OK;
    j->t_done = now_us();
    ++s->done;
    s->done_cv.notify_all();
    return MRT_OK;
}

void worker_main(mrt_server* s) {
    (void)hipSetDevice(s->ctx->device);
    uint32_t snap = 0;
    bool pending = false;  // a snapshot is in flight
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        int rc = retire(s);
        if (rc == MRT_OK && !pending && s->running.empty() && s->queue.size() == 1u && !s->broken &&
            direct_fits(s, s->jobs[s->queue.front()].get())) {
            rc = run_direct(s, s->jobs[s->queue.front()].get(), lk);
            if (rc == MRT_OK) continue;
        }
        if (rc == MRT_OK) rc = admit(s);
        bool issued = false;
        if (rc == MRT_OK) rc = issue_batch(s, snap, &issued);
        if (rc != MRT_OK) {
            // A device error ends every job: the device state is unknown.
            s->broken = true;
            s->broken_msg = g_error;
            for (uint64_t id : s->running) {
                Job* j = s->jobs[id].get();
                if (j->phase == kRunning || j->phase == kCopying) fail_job(s, j, rc, g_error);
            }
            s->running.clear();
            while (!s->queue.empty()) {
                fail_job(s, s->jobs[s->queue.front()].get(), rc, g_error);
                s->queue.pop_front();
            }
            return;
        }
        if (!issued && !pending) {
            // Nothing on the device: sleep until a submission, or poll the
            // output copies still in flight.
            if (s->running.empty() && s->queue.empty()) {
                if (s->shutdown) return;
                s->cv.wait(lk);
            } else {
                s->cv.wait_for(lk, std::chrono::microseconds(200));
            }
            continue;
        }
        if (pending) {
            const uint32_t prev = snap ^ 1u;
            hipEvent_t ev[kWfMaxGroups];
            uint32_t nev = 0;
            for (uint32_t g = 0; g < s->G; ++g)
                if (s->snap_mask[prev] & (1u << g)) ev[nev++] = s->chk[prev][g];
            lk.unlock();
            hipError_t q = hipSuccess;
            for (uint32_t k = 0; k < nev && q == hipSuccess; ++k) q = hipEventSynchronize(ev[k]);
            lk.lock();
            if (q != hipSuccess) {
                fail(MRT_E_HIP, std::string("hipEventSynchronize: ") + hipGetErrorString(q));
                rc = MRT_E_HIP;
            } else {
                rc = process_snapshot(s, prev);
            }
            if (rc != MRT_OK) {
                s->broken = true;
                s->broken_msg = g_error;
                for (uint64_t id : s->running) fail_job(s, s->jobs[id].get(), rc, g_error);
                s->running.clear();
                return;
            }
        }
        pending = issued;
        if (issued) snap ^= 1u;
    }
}

void fill_info(const Job* j, mrt_job_info* info) {
    if (!info) return;
    info->state = j->phase == kQueued ? MRT_JOB_QUEUED
                  : j->phase == kDone  ? MRT_JOB_DONE
                  : j->phase == kFailed ? MRT_JOB_FAILED
                                        : MRT_JOB_RUNNING;
    info->status = j->status;
    info->t_submit_us = j->t_submit;
    info->t_admit_us = j->t_admit;
    info->t_gpu_done_us = j->t_gpu;
    info->t_done_us = j->t_done;
    info->lanes = j->lanes;
    info->direct = j->direct ? 1u : 0u;
}

// Copies a resident scene's walk trees into the arena with rebased references
// (rebased on the device, on the transfer stream: the pipelines keep running,
// and the new nodes land past every node they read).
int add_to_arena(mrt_server* s, mrt_scene* sc, Resident* out) {
    const size_t nWide = sc->cw_nodes.bytes / (4 * sizeof(uint4));
    const size_t nWalk = sc->walk_nodes.bytes / sizeof(BNode);
    const size_t nGeo = sc->tri_geo.bytes / (3 * sizeof(float4));
    const size_t nRec = std::min(nGeo, sc->cw_recs.bytes / (8 * sizeof(float4)));
    if (s->geo_used + nGeo >= (1ull << kSlotBits) || s->wide_used + nWide >= 0xFFFFFFF0ull ||
        s->walk_used + nWalk >= 0xFFFFFFF0ull)
        return fail(MRT_E_BUDGET, "the scene arena's 27-bit triangle slots are used up");
    const uint32_t wideBase = static_cast<uint32_t>(s->wide_used);
    const uint32_t walkBase = static_cast<uint32_t>(s->walk_used);
    const uint32_t triBase = static_cast<uint32_t>(s->geo_used);
    const size_t cwBytes = (s->wide_used + nWide) * 4 * sizeof(uint4);
    const size_t recBytes = (s->geo_used + nGeo) * 8 * sizeof(float4);
    const size_t walkBytes = (s->walk_used + nWalk) * sizeof(BNode);
    const size_t geoBytes = (s->geo_used + nGeo) * 3 * sizeof(float4);
    const uint64_t total = cwBytes + recBytes + walkBytes + geoBytes;
    if (s->scene_memory && total > s->scene_memory)
        return fail(MRT_E_BUDGET, "scene arena needs " + std::to_string(total) + " bytes, limit " +
                                      std::to_string(s->scene_memory));
    int rc;
    auto cap = [](size_t need, size_t have) { return need <= have ? have : std::max(need, have + have / 2); };
    if ((rc = grow(s, s->arena_cw, cap(cwBytes, s->arena_cw.bytes), s->wide_used * 4 * sizeof(uint4))) != MRT_OK)
        return rc;
    if ((rc = grow(s, s->arena_recs, cap(recBytes, s->arena_recs.bytes), s->geo_used * 8 * sizeof(float4))) != MRT_OK)
        return rc;
    if ((rc = grow(s, s->arena_walk, cap(walkBytes, s->arena_walk.bytes), s->walk_used * sizeof(BNode))) != MRT_OK)
        return rc;
    if ((rc = grow(s, s->arena_geo, cap(geoBytes, s->arena_geo.bytes), s->geo_used * 3 * sizeof(float4))) != MRT_OK)
        return rc;
    // Compressed wide nodes: leaf words carry the slot in their low bits,
    // interior slots a wide node.  Walk pairs: leaf children (count != 0)
    // move with the slots, interior ones (finite box) with the pairs.  Leaf
    // records and triangle slots hold no references and move as they are.
    const hipStream_t st = s->xfer;
    MRT_HIP(gpu_rebase_cw(sc->cw_nodes.as<uint4>(), nWide, s->arena_cw.as<uint4>() + 4 * s->wide_used, triBase,
                          wideBase, st));
    MRT_HIP(gpu_rebase_walk(sc->walk_nodes.as<BNode>(), nWalk, s->arena_walk.as<BNode>() + s->walk_used, triBase,
                            walkBase, st));
    if (nGeo)
        MRT_HIP(hipMemcpyAsync(s->arena_geo.as<float4>() + 3 * s->geo_used, sc->tri_geo.p,
                               nGeo * 3 * sizeof(float4), hipMemcpyDeviceToDevice, st));
    if (nRec)
        MRT_HIP(hipMemcpyAsync(s->arena_recs.as<float4>() + 8 * s->geo_used, sc->cw_recs.p,
                               nRec * 8 * sizeof(float4), hipMemcpyDeviceToDevice, st));
    MRT_HIP(hipStreamSynchronize(st));
    s->wide_used += nWide;
    s->walk_used += nWalk;
    s->geo_used += nGeo;
    *out = Resident{sc, wideBase, walkBase, triBase};
    return MRT_OK;
}

void destroy_server(mrt_server* s) {
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->ctx->device);
    for (uint32_t g = 0; g < s->G; ++g)
        if (s->pipe[g]) (void)hipStreamSynchronize(s->pipe[g]);  // the context's streams
    if (s->xfer) {
        (void)hipStreamSynchronize(s->xfer);
        (void)hipStreamDestroy(s->xfer);
    }
    for (auto& row : s->chk)
        for (hipEvent_t e : row)
            if (e) (void)hipEventDestroy(e);
    for (auto& kv : s->jobs)
        if (kv.second->done_ev) (void)hipEventDestroy(kv.second->done_ev);
    if (s->jobs_host) (void)hipHostFree(s->jobs_host);
    if (s->snap_host) (void)hipHostFree(s->snap_host);
    if (s->staging) (void)hipHostFree(s->staging);
    {
        std::lock_guard<std::mutex> lock(s->ctx->mu);
        for (const Resident& r : s->scenes) --r.scene->servers;
    }
    --s->ctx->children;
    delete s;
}

}  // namespace

extern "C" {

MRT_API int mrt_server_create(mrt_ctx* ctx, const mrt_server_desc* desc, mrt_server** out) {
    if (!ctx || !out) return fail(MRT_E_INVALID, "null argument");
    *out = nullptr;
    const mrt_server_desc d = desc ? *desc : mrt_server_desc{};
    const uint32_t G = d.pipelines ? d.pipelines : kWfMaxGroups;
    if (G > kWfMaxGroups) return fail(MRT_E_INVALID, "at most 3 pipelines");
    const uint32_t slots = d.max_slots ? d.max_slots : kDefaultSlots;
    if (slots % 64u != 0u || slots / 64u < G) return fail(MRT_E_INVALID, "max_slots must be a multiple of 64 per pipeline");
    const uint32_t maxJobs = d.max_jobs ? d.max_jobs : kDefaultJobs;
    const uint32_t maxScenes = d.max_scenes ? d.max_scenes : kDefaultScenes;
    if (d.idle_direct > 1u) return fail(MRT_E_INVALID, "idle_direct must be 0 or 1");
    const uint64_t jobMem = align_up(d.job_memory ? d.job_memory : kDefaultJobMemory);
    const uint64_t sampleMem = align_up(d.sample_memory ? d.sample_memory : kDefaultSampleMemory);
    if (maxJobs > 65536u || maxScenes > 65536u) return fail(MRT_E_INVALID, "max_jobs and max_scenes must be <= 65536");
    const uint64_t need = wavefront_state_bytes(slots) + G * (wavefront_group_bytes(slots / G + 64u) + 256) +
                          jobMem + sampleMem + maxJobs * (sizeof(JobDev) + 4ull) + maxScenes * sizeof(SceneDev) +
                          (slots / 64ull) * 8ull;
    if (ctx->budget && need > ctx->budget)
        return fail(MRT_E_BUDGET, "render server needs " + std::to_string(need) + " bytes, budget " +
                                      std::to_string(ctx->budget));
    MRT_HIP(hipSetDevice(ctx->device));
    auto* s = new mrt_server;
    s->ctx = ctx;
    ++ctx->children;  // destroy_server gives it back
    s->G = G;
    s->max_jobs = maxJobs;
    s->max_scenes = maxScenes;
    s->pool_waves = slots / 64u;
    s->idle_direct = d.idle_direct == 1u;
    s->job_memory = jobMem;
    s->scene_memory = d.scene_memory;
    auto bail = [&](int rc) {
        destroy_server(s);
        return rc;
    };
    int rc = MRT_OK;
    {
        std::lock_guard<std::mutex> lock(ctx->mu);
        for (uint32_t g = 0; g < G; ++g)
            if ((rc = group_stream(ctx, g, &s->pipe[g])) != MRT_OK) return bail(rc);
    }
    for (uint32_t g = 0; g < G; ++g)
        for (uint32_t k = 0; k < 2; ++k)
            if (hipEventCreateWithFlags(&s->chk[k][g], hipEventDisableTiming) != hipSuccess)
                return bail(fail(MRT_E_HIP, "event creation failed"));
    if (hipStreamCreateWithFlags(&s->xfer, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(MRT_E_HIP, "stream creation failed"));
    if ((rc = s->state.alloc(wavefront_state_bytes(slots))) != MRT_OK) return bail(rc);
    float4* state = s->state.as<float4>();
    const uint32_t ncu = ctx->num_cu ? ctx->num_cu : 256u;
    s->trace_blocks = ncu * 16u / 2u;
    s->persist_blocks = wavefront_persistent_blocks(ncu, s->trace_blocks, G, ctx->tuning.launch_share_pct);
    for (uint32_t g = 0; g < G; ++g) {
        const uint32_t w0 = s->pool_waves * g / G, w1 = s->pool_waves * (g + 1) / G;
        WfBuffers& wb = s->wb[g];
        wb.slot_base = 64u * w0;
        wb.slot_count = 64u * (w1 - w0);
        const uint32_t n = wb.slot_count;
        if ((rc = s->group[g].alloc(wavefront_group_bytes(n) + 256)) != MRT_OK) return bail(rc);
        float4* base = s->group[g].as<float4>();
        wb.sA = state;
        wb.sB = state + slots;
        wb.sC = state + 2ull * slots;
        wb.queue_a = base;
        wb.queue_b = base + 2ull * n;
        wb.hits = base + 4ull * n;
        wb.shadow = base + 5ull * n;
        uint32_t* u = reinterpret_cast<uint32_t*>(base + 9ull * n);
        wb.counts = u;
        wb.ended[0] = u + 64;
        wb.ended[1] = wb.ended[0] + n;
        wb.keys = nullptr;
        wb.hist = nullptr;
        wb.offs = nullptr;
        wb.scan_temp = nullptr;
        wb.scan_bytes = 0;
        wb.sort = 0u;
        wb.refill_lanes = ctx->tuning.refill_lanes ? ctx->tuning.refill_lanes : kDefaultRefillLanes;
        wb.fetch_chunk = ctx->tuning.fetch_chunk ? ctx->tuning.fetch_chunk : kDefaultFetchChunk;
        MRT_HIP(hipMemsetAsync(wb.counts, 0, kCntWords * sizeof(uint32_t), s->pipe[g]));
        s->L[g].qin = wb.queue_a;
        s->L[g].qout = wb.queue_b;
        s->L[g].cin = kCntQueue0;
        s->L[g].cout = kCntQueue1;
        s->L[g].iter = 0;
        s->waves_free[g].reset(w0, w1 - w0);
    }
    if ((rc = s->jobs_dev.alloc(maxJobs * sizeof(JobDev))) != MRT_OK) return bail(rc);
    if ((rc = s->scenes_dev.alloc(maxScenes * sizeof(SceneDev))) != MRT_OK) return bail(rc);
    if ((rc = s->waves_dev.alloc(s->pool_waves * sizeof(uint2))) != MRT_OK) return bail(rc);
    if ((rc = s->finished_dev.alloc(G * maxJobs * sizeof(uint32_t))) != MRT_OK) return bail(rc);
    if ((rc = s->counters.alloc(16 * sizeof(unsigned long long))) != MRT_OK) return bail(rc);
    if ((rc = s->err.alloc(sizeof(uint32_t))) != MRT_OK) return bail(rc);
    if ((rc = s->job_mem.alloc(jobMem)) != MRT_OK) return bail(rc);
    if ((rc = s->sample_mem.alloc(sampleMem)) != MRT_OK) return bail(rc);
    MRT_HIP(hipMemset(s->counters.p, 0, 16 * sizeof(unsigned long long)));
    MRT_HIP(hipMemset(s->err.p, 0, sizeof(uint32_t)));
    MRT_HIP(hipMemset(s->finished_dev.p, 0, G * maxJobs * sizeof(uint32_t)));
    if (hipHostMalloc(reinterpret_cast<void**>(&s->jobs_host), maxJobs * sizeof(JobDev)) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&s->snap_host), 2ull * G * (maxJobs + 1u) * sizeof(uint32_t)) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&s->staging), jobMem) != hipSuccess)
        return bail(fail(MRT_E_NOMEM, "pinned host allocation failed"));
    std::memset(s->jobs_host, 0, maxJobs * sizeof(JobDev));
    s->mem_free.reset(0, jobMem / kAlign);
    s->sample_free.reset(0, sampleMem / kAlign);
    for (uint32_t i = 0; i < maxJobs; ++i) s->free_slots.push_back(maxJobs - 1u - i);
    if ((rc = ensure_spill(s, 0u)) != MRT_OK) return bail(rc);
    for (uint32_t g = 0; g < G; ++g) MRT_HIP(hipStreamSynchronize(s->pipe[g]));
    s->worker = std::thread(worker_main, s);
    *out = s;
    return MRT_OK;
}

MRT_API int mrt_server_destroy(mrt_server* s) {
    if (!s) return MRT_OK;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        s->shutdown = true;
        while (!s->queue.empty()) {
            Job* j = s->jobs[s->queue.front()].get();
            fail_job(s, j, MRT_E_STOPPED, "render server destroyed");
            s->queue.pop_front();
        }
        s->cv.notify_all();
    }
    destroy_server(s);
    return MRT_OK;
}

MRT_API int mrt_server_add_scene(mrt_server* s, mrt_scene* scene, uint32_t* scene_id) {
    if (!s || !scene || !scene_id) return fail(MRT_E_INVALID, "null argument");
    if (scene->ctx != s->ctx) return fail(MRT_E_INVALID, "scene belongs to another context");
    if (!scene->built || scene->generic || scene->dev.acc[MRT_CAT_TRIANGLE].kind != MRT_BVH ||
        !scene->dev.walk_nodes || !scene->dev.cw_nodes)
        return fail(MRT_E_UNSUPPORTED, "render servers take scenes of triangles and lights built with the BVH");
    std::lock_guard<std::mutex> lock(s->mu);
    if (s->scenes.size() >= s->max_scenes) return fail(MRT_E_BUDGET, "render server scene table is full");
    MRT_HIP(hipSetDevice(s->ctx->device));
    int rc;
    if ((rc = ensure_spill(s, std::max(s->depth_need, scene->depth_need))) != MRT_OK) return rc;
    s->depth_need = std::max(s->depth_need, scene->depth_need);
    Resident r{};
    if ((rc = add_to_arena(s, scene, &r)) != MRT_OK) return rc;
    const uint32_t id = static_cast<uint32_t>(s->scenes.size());
    MRT_HIP(hipMemcpy(s->scenes_dev.as<SceneDev>() + id, &scene->dev, sizeof(SceneDev), hipMemcpyHostToDevice));
    s->scenes.push_back(r);
    {
        std::lock_guard<std::mutex> clock(s->ctx->mu);
        ++scene->servers;
    }
    *scene_id = id;
    return MRT_OK;
}

MRT_API int mrt_server_submit(mrt_server* s, uint32_t scene_id, const mrt_render_desc* d, uint32_t* bitmap,
                              uint32_t stride_bytes, float* accum, int mem, uint64_t* job_id) {
    if (!s || !d || !bitmap || !job_id) return fail(MRT_E_INVALID, "null argument");
    if (mem != MRT_MEM_HOST && mem != MRT_MEM_DEVICE) return fail(MRT_E_INVALID, "mem must be host or device");
    auto j = std::make_unique<Job>();
    mrt_scene* scene;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        if (scene_id >= s->scenes.size()) return fail(MRT_E_INVALID, "unknown scene id");
        scene = s->scenes[scene_id].scene;
    }
    int rc;
    if ((rc = render_setup(scene, d, stride_bytes, &j->p, &j->rank)) != MRT_OK) return rc;
    if (!wavefront_eligible(scene, d) || d->kernel == MRT_KERNEL_MEGAKERNEL)
        return fail(MRT_E_UNSUPPORTED, "render servers run the wavefront pipeline: the PathTracer with "
                                       "samples_light 1, no reflective or transmissive materials, "
                                       "ray_depth_max >= 1 and at most 65535 samples per pixel");
    if (d->first_sample >= d->samples_pixel) return fail(MRT_E_INVALID, "first_sample must be below samples_pixel");
    for (uint32_t k = 0; k < 8; ++k)
        if (d->tile_skip[k] != 0u) return fail(MRT_E_UNSUPPORTED, "render servers render whole frames (tile_skip)");
    j->scene = scene_id;
    j->desc = *d;
    j->stride = stride_bytes;
    j->bitmap = bitmap;
    j->accum = accum;
    j->mem = mem;
    j->bm_bytes = static_cast<size_t>(d->height) * j->p.real_width * sizeof(uint32_t);
    j->acc_bytes = static_cast<size_t>(d->width) * d->height * 3 * sizeof(float);
    j->waves_px = wavefront_slots(j->p) / 64u;
    j->waves = j->waves_px;
    j->live = 256u * j->p.bsx * j->p.bsy;
    uint64_t bytes = align_up(kBlocks * sizeof(uint32_t));
    if (mem == MRT_MEM_HOST) {
        j->off_bm = bytes;
        bytes += align_up(j->bm_bytes);
        if (accum) {
            j->off_acc = bytes;
            bytes += align_up(j->acc_bytes);
        }
    }
    j->mem_bytes = bytes;
    std::lock_guard<std::mutex> lock(s->mu);
    if (s->broken) return fail(MRT_E_HIP, "render server stopped after an error: " + s->broken_msg);
    if (s->shutdown) return fail(MRT_E_INVALID, "render server is shutting down");
    if (j->waves_px > s->pool_waves) return fail(MRT_E_BUDGET, "frame needs more path slots than the server has");
    if (bytes > s->job_memory) return fail(MRT_E_BUDGET, "frame outputs exceed the server's job memory");
    j->id = s->next_id++;
    j->t_submit = now_us();
    *job_id = j->id;
    s->queue.push_back(j->id);
    s->jobs[j->id] = std::move(j);
    ++s->submitted;
    s->cv.notify_all();
    return MRT_OK;
}

MRT_API int mrt_server_poll(mrt_server* s, uint64_t job_id, mrt_job_info* info) {
    if (!s) return fail(MRT_E_INVALID, "null server");
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) return fail(MRT_E_INVALID, "unknown job id");
    fill_info(it->second.get(), info);
    return MRT_OK;
}

MRT_API int mrt_server_wait(mrt_server* s, uint64_t job_id, int64_t timeout_us, mrt_job_info* info) {
    if (!s) return fail(MRT_E_INVALID, "null server");
    std::unique_lock<std::mutex> lock(s->mu);
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) return fail(MRT_E_INVALID, "unknown job id");
    Job* j = it->second.get();
    auto ended = [&] { return j->phase == kDone || j->phase == kFailed; };
    if (timeout_us < 0) {
        s->done_cv.wait(lock, ended);
    } else if (!s->done_cv.wait_for(lock, std::chrono::microseconds(timeout_us), ended)) {
        fill_info(j, info);
        return fail(MRT_E_TIMEOUT, "job not done yet");
    }
    fill_info(j, info);
    const int status = j->status;
    const std::string msg = j->error;
    s->jobs.erase(it);
    if (status != MRT_OK) return fail(status, msg.empty() ? "render stopped" : msg);
    return MRT_OK;
}

MRT_API int mrt_server_cancel(mrt_server* s, uint64_t job_id) {
    if (!s) return fail(MRT_E_INVALID, "null server");
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) return fail(MRT_E_INVALID, "unknown job id");
    Job* j = it->second.get();
    if (j->phase == kQueued) {
        s->queue.erase(std::find(s->queue.begin(), s->queue.end(), job_id));
        fail_job(s, j, MRT_E_STOPPED, "render stopped");
        return MRT_OK;
    }
    if (j->direct) {
        // A one-launch render: it completes, then fails with MRT_E_STOPPED.
        if (j->phase == kRunning) j->cancel = true;
        return MRT_OK;
    }
    if (j->phase == kRunning && !j->cancel) {
        j->cancel = true;
        s->jobs_host[j->slot].stop = 1u;
        MRT_HIP(hipSetDevice(s->ctx->device));
        // On the job's pipelines, after their admit kernels wrote the entry.
        for (uint32_t g = 0; g < s->G; ++g)
            if (j->pipes & (1u << g))
                MRT_HIP(hipMemcpyAsync(&s->jobs_dev.as<JobDev>()[j->slot].stop, &s->jobs_host[j->slot].stop,
                                       sizeof(uint32_t), hipMemcpyHostToDevice, s->pipe[g]));
    }
    return MRT_OK;
}

MRT_API int mrt_server_stats(mrt_server* s, mrt_serving_stats* st) {
    if (!s || !st) return fail(MRT_E_INVALID, "null argument");
    unsigned long long cnt[16] = {0};
    {
        std::lock_guard<std::mutex> lock(s->mu);
        MRT_HIP(hipSetDevice(s->ctx->device));
        std::memset(st, 0, sizeof(*st));
        st->jobs_submitted = s->submitted;
        st->jobs_done = s->done;
        st->jobs_failed = s->failed;
        st->jobs_queued = static_cast<uint32_t>(s->queue.size());
        st->jobs_running = static_cast<uint32_t>(s->running.size()) + (s->direct_job ? 1u : 0u);
        st->slots_total = s->pool_waves * 64u;
        uint64_t freeWaves = 0;
        for (uint32_t g = 0; g < s->G; ++g) freeWaves += s->waves_free[g].free_total();
        st->slots_used = static_cast<uint32_t>((s->pool_waves - freeWaves) * 64u);
        st->scene_bytes = s->wide_used * 4 * sizeof(uint4) + s->geo_used * 8 * sizeof(float4) +
                          s->walk_used * sizeof(BNode) +
                          s->geo_used * 3 * sizeof(float4);
        st->job_bytes_used = (s->mem_free.total() - s->mem_free.free_total()) * kAlign;
        st->batches = s->batches;
    }
    // The counters are read while kernels may be running: a snapshot.
    uint64_t direct[3];
    {
        std::lock_guard<std::mutex> lock(s->mu);
        std::copy(s->direct_rays, s->direct_rays + 3, direct);
    }
    MRT_HIP(hipMemcpy(cnt, s->counters.p, sizeof(cnt), hipMemcpyDeviceToHost));
    st->rays_camera = cnt[0] + direct[0];
    st->rays_extension = cnt[1] + direct[1];
    st->rays_shadow = cnt[2] + direct[2];
    return MRT_OK;
}

}  // extern "C"

// Render service: many jobs on one device (dl_hip_service_* in dl_hip.h).
//
// Lanes are worker threads, each owning one backend context (its own HIP
// stream, scene, BVH and render state); kernels of jobs on different lanes
// overlap on the device's hardware queues.  Within a lane, jobs share
// wavefront passes (dl_hip_render_batch): a lane that starts a job keeps
// taking further queued jobs with compatible settings into its running group
// at every pass boundary, while the pass stays within the pass size, so small
// jobs travel together through one set of queues and kernels (continuous
// batching).  A joining job's scene is the group's or any scene resident on
// the lane: the paths of a pass over several scenes carry their scene's index
// into one device layout holding the lane's resident scenes (unless
// single_scene_passes asks for one scene per pass).  Each job's passes have
// the sizes its standalone render would use, so its image is still
// bit-identical to a single dl_hip_render call.  A lane keeps every scene it has uploaded resident
// (dl_hip_scene_keep / dl_hip_scene_select: switching scenes costs no upload
// and no BVH build), and prefers a queued job of one of its resident scenes
// among the next `lanes` queue entries (bounded bypass, so no job waits
// behind more than `lanes` later arrivals for that reason).
//
// Admission: every lane holds a reservation of device bytes -- while it runs
// a job, the larger of what its context already holds and a conservative
// estimate of what it can reach during the job (fixed context buffers +
// its other resident scenes, as measured + the job's scene and BVH build
// scratch + render state); when idle, the bytes its context actually
// holds.  Under budget pressure idle lanes free everything they hold.  A job starts on a lane only if the
// sum of all reservations stays within the budget, so the sum of the lanes'
// device memory never exceeds it.  After each job the lane's measured peak
// is compared with its reservation (estimate_misses counts violations).
//
// A job renders in passes of its spp_per_pass, or of at most kPathsPerPass
// paths when that is 0 (half dl_hip_render's own default, so that smaller
// jobs can share a pass), accumulating into its film slot (or, for a guided
// job, which renders alone, the context's film), so its image is
// bit-identical to a single dl_hip_render call with that spp_per_pass;
// cancellation takes effect between passes.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/dl_hip.h"
#include "../render_sizes.h"

namespace {

using Clock = std::chrono::steady_clock;

constexpr uint64_t kPathsPerPass = 64u << 20;  // paths per pass of a job or a shared group (half the largest pass)
constexpr uint64_t kStateBytesPerPath = dl_sizes::kStateBytesPerPath;  // path state + queues
constexpr uint64_t kFilmBytesPerPixel = dl_sizes::kFilmBytesPerPixel;  // accumulation film + resolved RGB
constexpr uint64_t kBytesPerVertex = 44;        // positions, normals, tangents, texture coordinates
constexpr uint64_t kBytesPerPrim = 16 + 112 + 256;  // indices + mesh map, BVH nodes + triangles, build scratch
constexpr uint64_t kBytesPerTexel = 32;         // environment texels + alias tables
constexpr uint64_t kSlackBytes = 4u << 20;

// A registered scene: deep copies of every array the descriptor points to.
struct SceneEntry {
    std::vector<std::vector<float>> fbuf;
    std::vector<std::vector<uint32_t>> ubuf;
    std::vector<dl_mesh_desc> meshes;
    std::vector<dl_material_desc> materials;
    std::vector<dl_light_desc> lights;
    std::vector<dl_texture_desc> textures;
    dl_scene_desc desc{};
    uint64_t verts = 0, prims = 0, texels = 0;
    uint64_t bytes_estimate() const {
        return verts * kBytesPerVertex + prims * kBytesPerPrim + texels * kBytesPerTexel + kSlackBytes;
    }
};

struct Job {
    uint64_t id = 0, scene = 0;
    dl_camera_desc camera{};  // the scene's camera unless the job gave one
    dl_render_params params{};
    dl_job_info info{};
    std::vector<float> rgb;
    std::string error;
    std::atomic<bool> cancel{false};  // set by dl_hip_service_cancel / stop, read between passes
};

struct Lane {
    int index = 0;
    void* ctx = nullptr;
    uint64_t fixed_bytes = 0;     // context buffers that exist without a scene
    uint64_t resident_scene = 0;  // the context's current scene (0: none)
    std::map<uint64_t, uint64_t> kept;  // resident scenes -> device bytes they hold (dl_hip_scene_bytes)
    uint64_t state_paths = 0, state_pixels = 0;  // render state the context holds (paths, film pixels)
    uint64_t slot_pixels[DL_BATCH_MAX] = {};     // film slots the context holds (pixels each)
    uint64_t reservation = 0;
    std::thread th;
};

struct Service {
    dl_service_config cfg{};
    Clock::time_point t0;
    std::mutex m;
    std::condition_variable cv;
    bool stopping = false;
    bool pressure = false;  // a queued job waits for budget: idle lanes release their resident bytes
    uint64_t next_scene = 1, next_job = 1;
    std::map<uint64_t, std::unique_ptr<SceneEntry>> scenes;
    std::map<uint64_t, std::unique_ptr<Job>> jobs;
    std::deque<uint64_t> queue;
    std::vector<std::unique_ptr<Lane>> lanes;
    dl_service_stats stats{};
    std::string err;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
    uint64_t reserved_total() const {
        uint64_t s = 0;
        for (const auto& l : lanes) s += l->reservation;
        return s;
    }
};

uint64_t job_pixels(const Job& j) { return (uint64_t)j.camera.width * j.camera.height; }

// samples per pixel per pass of the job
uint32_t pass_spp(const dl_render_params& p, uint64_t pixels) {
    uint64_t k = p.spp_per_pass > 0 ? (uint64_t)p.spp_per_pass : std::max<uint64_t>(1, kPathsPerPass / pixels);
    return (uint32_t)std::min<uint64_t>(k, (uint64_t)p.spp);
}

uint64_t union_reservation(const Service& S, const Lane& lane, uint64_t scene_id, const SceneEntry& s);

// Device bytes the lane may reach while rendering `job` (see file comment;
// a layout of several scenes that an earlier group left is still held).
uint64_t reservation_for(const Service& S, const Lane& lane, const Job& job, const SceneEntry& s) {
    const uint64_t px = job_pixels(job);
    const uint64_t paths = px * pass_spp(job.params, px);
    uint64_t state = std::max(paths, lane.state_paths) * kStateBytesPerPath +
                     std::max(px, lane.state_pixels) * kFilmBytesPerPixel + kSlackBytes;
    // a guided render adds its mixtures, training records, network and
    // training batch
    if (job.params.flags & DL_RENDER_GUIDED) state += dl_sizes::guide_bytes(std::max(paths, lane.state_paths));
    uint64_t others = 0;
    for (const auto& kv : lane.kept)
        if (kv.first != job.scene) others += kv.second;
    const auto it = lane.kept.find(job.scene);
    const uint64_t scene = it != lane.kept.end() ? it->second : s.bytes_estimate();
    return lane.fixed_bytes + others + scene + state + union_reservation(S, lane, job.scene, s);
}

// Device bytes of the layout that holds every resident scene of the lane in
// one (built by the first pass over several scenes, kept while they stay):
// about the bytes those scenes hold.  Counted once the lane holds two scenes,
// as a pass over several of them may build it at any time.
uint64_t union_reservation(const Service& S, const Lane& lane, uint64_t scene_id, const SceneEntry& s) {
    if (S.cfg.single_scene_passes) return 0;
    uint64_t sum = 0, n = 0;
    for (const auto& kv : lane.kept) {
        sum += kv.second;
        n++;
    }
    if (!lane.kept.count(scene_id)) {
        sum += s.bytes_estimate();
        n++;
    }
    return n >= 2 ? sum + kSlackBytes : 0;
}

// Device bytes the lane may reach while running a group of shared passes of
// at most `pass_paths` paths led by a job of scene `scene_id`, whose film
// slots need slot_px[s] pixels (see file comment).
uint64_t group_reservation(const Service& S, const Lane& lane, uint64_t scene_id, const SceneEntry& s,
                           uint64_t pass_paths, const uint64_t* slot_px) {
    uint64_t films = 0;
    for (int k = 0; k < DL_BATCH_MAX; k++) films += std::max(lane.slot_pixels[k], slot_px[k]);
    const uint64_t state = std::max(pass_paths, lane.state_paths) * kStateBytesPerPath +
                           (lane.state_pixels + films) * kFilmBytesPerPixel + kSlackBytes;
    uint64_t others = 0;
    for (const auto& kv : lane.kept)
        if (kv.first != scene_id) others += kv.second;
    const auto it = lane.kept.find(scene_id);
    const uint64_t scene = it != lane.kept.end() ? it->second : s.bytes_estimate();
    return lane.fixed_bytes + others + scene + state + union_reservation(S, lane, scene_id, s);
}

// Jobs that can share a pass: the settings a pass shares, and either the
// same scene or (passes over several scenes) b's scene resident on the lane
// that runs a's group.
bool can_share(const Service& S, const Lane& lane, const Job& a, const Job& b) {
    const bool scene_ok = a.scene == b.scene || (!S.cfg.single_scene_passes && lane.kept.count(b.scene) > 0);
    return scene_ok && a.params.max_depth == b.params.max_depth && a.params.rr_depth == b.params.rr_depth &&
           !(a.params.flags & DL_RENDER_GUIDED) && !(b.params.flags & DL_RENDER_GUIDED) &&
           (a.params.flags & DL_RENDER_NORMALS) == (b.params.flags & DL_RENDER_NORMALS);
}

// Pick the next job for `lane` (caller holds the lock): policy order, with a
// preference for the lane's resident scene among the next `lanes` entries.
std::deque<uint64_t>::iterator pick(Service& S, Lane& lane) {
    auto& q = S.queue;
    if (q.empty()) return q.end();
    auto best = q.begin();
    if (S.cfg.policy == DL_POLICY_SJF) {
        uint64_t bw = ~0ull;
        for (auto it = q.begin(); it != q.end(); ++it) {
            const Job& j = *S.jobs[*it];
            const uint64_t w = job_pixels(j) * (uint64_t)j.params.spp;
            if (w < bw) {
                bw = w;
                best = it;
            }
        }
        return best;
    }
    const size_t window = std::min<size_t>(q.size(), (size_t)S.cfg.lanes);
    for (size_t i = 0; i < window; i++)
        if (S.jobs[q[i]]->scene == lane.resident_scene) return q.begin() + (long)i;
    for (size_t i = 0; i < window; i++)
        if (lane.kept.count(S.jobs[q[i]]->scene)) return q.begin() + (long)i;
    return best;
}

void finish(Service& S, Job& job, int state, const std::string& error) {
    job.info.state = state;
    job.info.t_finish = S.now();
    job.error = error;
    if (state == DL_JOB_DONE) S.stats.jobs_done++;
    if (state == DL_JOB_FAILED) S.stats.jobs_failed++;
    if (state == DL_JOB_CANCELLED) S.stats.jobs_cancelled++;
}

// Make the job's scene current on the lane's context (no lock held).
// switch: 0 the scene is current, 1 select a resident scene, 2 upload and
// build (then keep it resident; *kept_bytes = the device bytes it added).
std::string prepare_scene(Service& S, Lane& lane, const Job& job, const SceneEntry& scene, int switch_,
                          uint64_t* kept_bytes) {
    void* c = lane.ctx;
    auto fail = [&]() { return std::string(dl_hip_last_error(c)); };
    if (switch_ == 1 && dl_hip_scene_select(c, job.scene) != DL_OK) return fail();
    if (switch_ == 2) {
        // the current scene stays resident (it was kept)
        if (dl_hip_upload_scene(c, &scene.desc) != DL_OK) return fail();
        if (dl_hip_build_accel(c, S.cfg.builder, nullptr) != DL_OK) return fail();
        if (dl_hip_scene_keep(c, job.scene) != DL_OK) return fail();
        if (dl_hip_scene_bytes(c, job.scene, kept_bytes) != DL_OK) return fail();
    }
    (void)S;
    return "";
}

// Render one job alone on the lane's context, its scene current (no lock
// held).  Returns "" or an error.
std::string run_job(Service& S, Lane& lane, Job& job, dl_render_stats& total) {
    (void)S;
    void* c = lane.ctx;
    auto fail = [&]() { return std::string(dl_hip_last_error(c)); };
    if (dl_hip_set_camera(c, &job.camera) != DL_OK) return fail();
    const uint64_t px = job_pixels(job);
    const uint32_t k = pass_spp(job.params, px);
    job.rgb.assign(px * 3, 0.0f);
    std::memset(&total, 0, sizeof(total));
    for (uint32_t done = 0; done < (uint32_t)job.params.spp;) {
        if (job.cancel.load()) return "cancelled";
        const uint32_t kk = std::min<uint32_t>(k, (uint32_t)job.params.spp - done);
        dl_render_params p = job.params;
        p.spp = (int32_t)kk;
        p.spp_per_pass = (int32_t)k;
        p.first_sample = job.params.first_sample + (int32_t)done;
        p.flags = (job.params.flags & ~DL_RENDER_ACCUMULATE) | (done > 0 ? DL_RENDER_ACCUMULATE : 0);
        const bool last = done + kk >= (uint32_t)job.params.spp;
        dl_render_stats st{};
        if (dl_hip_render(c, &p, last ? job.rgb.data() : nullptr, &st) != DL_OK) return fail();
        total.paths += st.paths;
        total.ms_total += st.ms_total;
        done += kk;
    }
    lane.state_paths = std::max<uint64_t>(lane.state_paths, px * k);
    lane.state_pixels = std::max<uint64_t>(lane.state_pixels, px);
    return "";
}

// A group of shared passes led by `lead` (its scene current, unless
// `prep_error` says why not; no lock held on entry or exit).  At every pass
// boundary the group drops finished and cancelled jobs and takes queued jobs
// that can share its passes while the pass stays within kPathsPerPass paths,
// the group within DL_BATCH_MAX jobs and the lane's reservation within the
// budget.
struct Member {
    Job* job;
    uint32_t k = 0, done = 0;  // pass size and samples rendered
    int slot = 0;
    double paths = 0, ms = 0;
};

void run_group(Service& S, Lane& lane, Job* lead, const SceneEntry& scene, std::string prep_error, int switch_,
               uint64_t kept_bytes) {
    void* c = lane.ctx;
    std::vector<Member> group;
    const uint64_t lpx = job_pixels(*lead);
    Member m0;
    m0.job = lead;
    m0.k = pass_spp(lead->params, lpx);
    group.push_back(m0);
    lead->rgb.assign(lpx * 3, 0.0f);
    std::string error = prep_error;
    uint64_t peak_all = 0, max_pass = 0;
    bool ok = error.empty();
    std::unique_lock<std::mutex> lk(S.m);
    if (ok && switch_ == 2) lane.kept[lead->scene] = kept_bytes;
    if (ok) lane.resident_scene = lead->scene;
    while (!group.empty()) {
        // drop cancelled jobs
        for (size_t i = 0; i < group.size();) {
            if (group[i].job->cancel.load() || !ok) {
                Job& j = *group[i].job;
                j.info.paths = group[i].paths;
                j.info.ms_gpu = group[i].ms;
                S.stats.paths += group[i].paths;
                finish(S, j, ok ? DL_JOB_CANCELLED : DL_JOB_FAILED, ok ? "" : error);
                group.erase(group.begin() + (long)i);
                S.cv.notify_all();
            } else {
                ++i;
            }
        }
        if (group.empty()) break;
        // take queued jobs that can share the passes
        uint64_t pass_paths = 0;
        bool slot_used[DL_BATCH_MAX] = {};
        uint64_t slot_px[DL_BATCH_MAX] = {};
        for (const Member& g : group) {
            const uint64_t px = job_pixels(*g.job);
            pass_paths += px * std::min<uint32_t>(g.k, (uint32_t)g.job->params.spp - g.done);
            slot_used[g.slot] = true;
            slot_px[g.slot] = px;
        }
        const size_t max_group = S.cfg.max_batch > 0 ? (size_t)std::min(S.cfg.max_batch, (int)DL_BATCH_MAX)
                                                     : (size_t)DL_BATCH_MAX;
        for (auto qi = S.queue.begin(); qi != S.queue.end() && group.size() < max_group;) {
            Job& cand = *S.jobs[*qi];
            const uint64_t px = job_pixels(cand);
            const uint32_t k = pass_spp(cand.params, px);
            if (!can_share(S, lane, *lead, cand) || pass_paths + px * k > kPathsPerPass) {
                ++qi;
                continue;
            }
            int slot = 0;
            while (slot_used[slot]) ++slot;
            uint64_t try_px[DL_BATCH_MAX];
            std::copy(slot_px, slot_px + DL_BATCH_MAX, try_px);
            try_px[slot] = px;
            const uint64_t need =
                std::max(lane.reservation, group_reservation(S, lane, lead->scene, scene, pass_paths + px * k, try_px));
            if (S.reserved_total() - lane.reservation + need > S.cfg.budget_bytes) {
                ++qi;
                continue;
            }
            lane.reservation = need;
            S.stats.peak_reserved_bytes = std::max<uint64_t>(S.stats.peak_reserved_bytes, S.reserved_total());
            cand.info.state = DL_JOB_RUNNING;
            cand.info.lane = lane.index;
            cand.info.t_start = S.now();
            cand.info.reserved_bytes = need;
            cand.rgb.assign(px * 3, 0.0f);
            Member mm;
            mm.job = &cand;
            mm.k = k;
            mm.slot = slot;
            group.push_back(mm);
            slot_used[slot] = true;
            slot_px[slot] = px;
            pass_paths += px * k;
            S.stats.batched_jobs++;
            qi = S.queue.erase(qi);
        }
        if (group.size() > 1) S.stats.shared_passes++;
        for (const Member& g : group)
            if (g.job->scene != lead->scene) {
                S.stats.cross_scene_passes++;
                break;
            }
        // every member runs under the lane's current reservation
        for (const Member& g : group)
            g.job->info.reserved_bytes = std::max<uint64_t>(g.job->info.reserved_bytes, lane.reservation);
        lk.unlock();

        std::vector<dl_batch_item> items(group.size());
        for (size_t i = 0; i < group.size(); i++) {
            const Member& g = group[i];
            const Job& j = *g.job;
            const uint32_t kk = std::min<uint32_t>(g.k, (uint32_t)j.params.spp - g.done);
            dl_batch_item& it = items[i];
            it.camera = j.camera;
            it.spp = (int32_t)kk;
            it.first_sample = j.params.first_sample + (int32_t)g.done;
            it.seed = j.params.seed;
            it.film = g.slot;
            it.accumulate = g.done > 0 ? 1 : 0;
            it.rgb_out = g.done + kk >= (uint32_t)j.params.spp ? g.job->rgb.data() : nullptr;
            it.scene = j.scene;  // every scene the lane uploads is kept under its service id
        }
        dl_render_params shared = lead->params;
        dl_render_stats st{};
        const bool pass_ok = dl_hip_render_batch(c, &shared, items.data(), (uint32_t)items.size(), &st) == DL_OK;
        if (!pass_ok) error = dl_hip_last_error(c);
        uint64_t cur = 0, peak = 0;
        dl_hip_memory_usage(c, &cur, &peak);
        peak_all = std::max(peak_all, peak);
        max_pass = std::max<uint64_t>(max_pass, (uint64_t)st.paths);

        lk.lock();
        if (!pass_ok) {
            ok = false;
            continue;  // the group fails at the top of the loop
        }
        // the render state keeps the capacity of the group's largest pass:
        // later reservations of the group start from it
        lane.state_paths = std::max<uint64_t>(lane.state_paths, (uint64_t)st.paths);
        for (size_t i = 0; i < group.size(); i++) {
            Member& g = group[i];
            const double p = (double)job_pixels(*g.job) * items[i].spp;
            g.paths += p;
            g.ms += st.paths > 0 ? st.ms_total * p / st.paths : 0.0;
            g.done += (uint32_t)items[i].spp;
            const uint64_t px = job_pixels(*g.job);
            lane.slot_pixels[g.slot] = std::max(lane.slot_pixels[g.slot], px);
            g.job->info.peak_bytes = std::max<uint64_t>(g.job->info.peak_bytes, peak);
        }
        for (size_t i = 0; i < group.size();) {
            Member& g = group[i];
            if (g.done >= (uint32_t)g.job->params.spp) {
                Job& j = *g.job;
                j.info.paths = g.paths;
                j.info.ms_gpu = g.ms;
                if (j.info.peak_bytes > j.info.reserved_bytes) S.stats.estimate_misses++;
                S.stats.paths += g.paths;
                finish(S, j, DL_JOB_DONE, "");
                group.erase(group.begin() + (long)i);
                S.cv.notify_all();
            } else {
                ++i;
            }
        }
    }
    lane.state_paths = std::max<uint64_t>(lane.state_paths, max_pass);
    if (!ok) {
        // the context may be left between scenes: start over from nothing
        lk.unlock();
        dl_hip_trim(c);
        lk.lock();
        lane.kept.clear();
        lane.resident_scene = 0;
        lane.state_paths = lane.state_pixels = 0;
        for (auto& v : lane.slot_pixels) v = 0;
    }
    lk.unlock();
    uint64_t cur = 0;
    dl_hip_memory_usage(c, &cur, nullptr);
    lk.lock();
    lane.reservation = cur;
    S.cv.notify_all();
    (void)peak_all;
}

void lane_main(Service* Sp, Lane* lp) {
    Service& S = *Sp;
    Lane& lane = *lp;
    std::unique_lock<std::mutex> lk(S.m);
    for (;;) {
        S.cv.wait(lk, [&] {
            return S.stopping || !S.queue.empty() || (S.pressure && lane.reservation > lane.fixed_bytes);
        });
        if (S.stopping) break;
        if (S.pressure && lane.reservation > lane.fixed_bytes) {
            // idle under budget pressure: drop the resident scene and state
            lk.unlock();
            dl_hip_trim(lane.ctx);
            uint64_t cur = 0;
            dl_hip_memory_usage(lane.ctx, &cur, nullptr);
            lk.lock();
            lane.resident_scene = 0;
            lane.kept.clear();
            lane.state_paths = lane.state_pixels = 0;
            for (auto& v : lane.slot_pixels) v = 0;
            lane.reservation = cur;
            S.cv.notify_all();
        }
        auto it = pick(S, lane);
        if (it == S.queue.end()) continue;
        Job& job = *S.jobs[*it];
        const SceneEntry& scene = *S.scenes[job.scene];
        const bool alone = (job.params.flags & DL_RENDER_GUIDED) != 0;  // guided jobs render alone
        const uint64_t px = job_pixels(job);
        uint64_t slot_px[DL_BATCH_MAX] = {};
        slot_px[0] = px;
        // the lane starts from what it holds (its previous scene stays until
        // the upload replaces it), then may reach the estimate
        const uint64_t est = alone ? reservation_for(S, lane, job, scene)
                                   : group_reservation(S, lane, job.scene, scene, px * pass_spp(job.params, px),
                                                       slot_px);
        const uint64_t need = std::max(lane.reservation, est);
        const uint64_t others = S.reserved_total() - lane.reservation;
        if (others + need > S.cfg.budget_bytes) {
            // with every other lane at its floor (fixed buffers only) the job
            // still would not fit: it never will
            uint64_t floor_others = 0;
            for (const auto& l : S.lanes)
                if (l.get() != &lane) floor_others += l->fixed_bytes;
            if (floor_others + need > S.cfg.budget_bytes) {
                S.queue.erase(it);
                finish(S, job, DL_JOB_FAILED, "job needs an estimated " + std::to_string(need) +
                                                  " device bytes; the budget is " + std::to_string(S.cfg.budget_bytes));
                S.cv.notify_all();
                continue;
            }
            // otherwise wait for a lane to finish or to release resident bytes
            S.pressure = true;
            S.cv.notify_all();
            S.cv.wait_for(lk, std::chrono::milliseconds(2));
            continue;
        }
        S.pressure = false;
        S.queue.erase(it);
        lane.reservation = need;
        S.stats.peak_reserved_bytes = std::max<uint64_t>(S.stats.peak_reserved_bytes, S.reserved_total());
        job.info.state = DL_JOB_RUNNING;
        job.info.lane = lane.index;
        job.info.t_start = S.now();
        job.info.reserved_bytes = need;
        const int switch_ = lane.resident_scene == job.scene ? 0 : lane.kept.count(job.scene) ? 1 : 2;
        if (switch_ == 1) S.stats.scene_selects++;
        if (switch_ == 2) S.stats.scene_uploads++;
        Job* jp = &job;
        lk.unlock();

        // the context's peak counter restarts at each query: the next query
        // gives the peak of this job (or group)
        dl_hip_memory_usage(lane.ctx, nullptr, nullptr);
        uint64_t kept_bytes = 0;
        std::string e = prepare_scene(S, lane, *jp, scene, switch_, &kept_bytes);
        if (alone) {
            dl_render_stats total{};
            if (e.empty()) e = run_job(S, lane, *jp, total);
            uint64_t cur = 0, peak = 0;
            dl_hip_memory_usage(lane.ctx, &cur, &peak);
            lk.lock();
            const bool ok = e.empty() || e == "cancelled";
            if (switch_ == 2 && ok) lane.kept[jp->scene] = kept_bytes;
            if (!ok) {
                // the context may be left between scenes: start over from nothing
                lk.unlock();
                dl_hip_trim(lane.ctx);
                dl_hip_memory_usage(lane.ctx, &cur, nullptr);
                lk.lock();
                lane.kept.clear();
                lane.state_paths = lane.state_pixels = 0;
                for (auto& v : lane.slot_pixels) v = 0;
            }
            lane.resident_scene = ok ? jp->scene : 0;
            jp->info.paths = total.paths;
            jp->info.ms_gpu = total.ms_total;
            jp->info.peak_bytes = peak;
            if (peak > need) S.stats.estimate_misses++;
            S.stats.paths += total.paths;
            if (e.empty())
                finish(S, *jp, DL_JOB_DONE, "");
            else if (e == "cancelled")
                finish(S, *jp, DL_JOB_CANCELLED, "");
            else
                finish(S, *jp, DL_JOB_FAILED, e);
            lane.reservation = cur;
            S.cv.notify_all();
            continue;
        }
        run_group(S, lane, jp, scene, e, switch_, kept_bytes);
        lk.lock();
    }
}

thread_local std::string g_service_error;

dl_status err(Service* S, dl_status s, const std::string& msg) {
    if (S)
        S->err = msg;
    else
        g_service_error = msg;
    return s;
}

}  // namespace

extern "C" {

dl_status dl_hip_service_start(const dl_service_config* cfg, void** svc) {
    if (!cfg || !svc) return err(nullptr, DL_ERR_INVALID, "null argument");
    *svc = nullptr;
    if (cfg->lanes < 1 || cfg->lanes > 16) return err(nullptr, DL_ERR_INVALID, "lanes must be in [1, 16]");
    if (cfg->policy != DL_POLICY_FIFO && cfg->policy != DL_POLICY_SJF)
        return err(nullptr, DL_ERR_INVALID, "unknown scheduling policy");
    if (cfg->max_batch < 0) return err(nullptr, DL_ERR_INVALID, "max_batch must be >= 0");
    std::unique_ptr<Service> S(new Service);
    S->cfg = *cfg;
    S->stats.budget_bytes = cfg->budget_bytes;
    for (int i = 0; i < cfg->lanes; i++) {
        std::unique_ptr<Lane> l(new Lane);
        l->index = i;
        const dl_status st = dl_hip_create(cfg->device, &l->ctx);
        if (st != DL_OK) {
            const std::string msg = l->ctx ? dl_hip_last_error(l->ctx) : "no HIP device";
            if (l->ctx) dl_hip_destroy(l->ctx);
            for (auto& o : S->lanes) dl_hip_destroy(o->ctx);
            return err(nullptr, st, msg);
        }
        dl_hip_memory_usage(l->ctx, &l->fixed_bytes, nullptr);
        l->reservation = l->fixed_bytes;
        S->lanes.push_back(std::move(l));
    }
    if (S->reserved_total() > cfg->budget_bytes) {
        for (auto& o : S->lanes) dl_hip_destroy(o->ctx);
        return err(nullptr, DL_ERR_INVALID, "the lanes' fixed buffers alone exceed the budget");
    }
    S->t0 = Clock::now();
    for (auto& l : S->lanes) l->th = std::thread(lane_main, S.get(), l.get());
    *svc = S.release();
    return DL_OK;
}

dl_status dl_hip_service_add_scene(void* h, const dl_scene_desc* d, uint64_t* scene_id) {
    Service* S = (Service*)h;
    if (!S || !d || !scene_id) return err(S, DL_ERR_INVALID, "null argument");
    if (d->camera.width <= 0 || d->camera.height <= 0) return err(S, DL_ERR_INVALID, "invalid camera size");
    std::unique_ptr<SceneEntry> e(new SceneEntry);
    e->desc = *d;
    e->fbuf.reserve(4 * d->num_meshes + d->num_lights + d->num_textures);
    e->ubuf.reserve(d->num_meshes);
    for (uint32_t i = 0; i < d->num_meshes; i++) {
        dl_mesh_desc m = d->meshes[i];
        if (!m.positions || !m.indices) return err(S, DL_ERR_INVALID, "mesh without positions or indices");
        auto copyf = [&](const float*& p, size_t n) {
            if (!p) return;
            e->fbuf.emplace_back(p, p + n);
            p = e->fbuf.back().data();
        };
        copyf(m.positions, 3 * (size_t)m.num_vertices);
        copyf(m.normals, 3 * (size_t)m.num_vertices);
        copyf(m.texcoords, 2 * (size_t)m.num_vertices);
        copyf(m.tangents, 3 * (size_t)m.num_vertices);
        e->ubuf.emplace_back(m.indices, m.indices + 3 * (size_t)m.num_triangles);
        m.indices = e->ubuf.back().data();
        e->verts += m.num_vertices;
        e->prims += m.num_triangles;
        e->meshes.push_back(m);
    }
    e->materials.assign(d->materials, d->materials + d->num_materials);
    for (uint32_t i = 0; i < d->num_lights; i++) {
        dl_light_desc l = d->lights[i];
        if (l.type == DL_LIGHT_ENV && l.env_rgb) {
            const size_t n = (size_t)l.env_width * l.env_height;
            e->fbuf.emplace_back(l.env_rgb, l.env_rgb + 3 * n);
            l.env_rgb = e->fbuf.back().data();
            e->texels += n;
        }
        e->lights.push_back(l);
    }
    for (uint32_t i = 0; i < d->num_textures; i++) {
        dl_texture_desc t = d->textures[i];
        if (!t.rgb || t.width <= 0 || t.height <= 0) return err(S, DL_ERR_INVALID, "invalid texture");
        const size_t n = (size_t)t.width * t.height;
        e->fbuf.emplace_back(t.rgb, t.rgb + 3 * n);
        t.rgb = e->fbuf.back().data();
        e->texels += n;
        e->textures.push_back(t);
    }
    e->desc.meshes = e->meshes.data();
    e->desc.materials = e->materials.data();
    e->desc.lights = e->lights.data();
    e->desc.textures = e->textures.data();
    std::lock_guard<std::mutex> lk(S->m);
    *scene_id = S->next_scene++;
    S->scenes[*scene_id] = std::move(e);
    return DL_OK;
}

dl_status dl_hip_service_submit(void* h, uint64_t scene_id, const dl_camera_desc* camera, const dl_render_params* p,
                                uint64_t* job_id) {
    Service* S = (Service*)h;
    if (!S || !p || !job_id) return err(S, DL_ERR_INVALID, "null argument");
    if (p->spp <= 0) return err(S, DL_ERR_INVALID, "spp must be positive");
    if (camera && (camera->width <= 0 || camera->height <= 0)) return err(S, DL_ERR_INVALID, "bad camera size");
    std::lock_guard<std::mutex> lk(S->m);
    if (S->stopping) return err(S, DL_ERR_STATE, "the service is stopping");
    if (!S->scenes.count(scene_id)) return err(S, DL_ERR_INVALID, "unknown scene id " + std::to_string(scene_id));

    std::unique_ptr<Job> j(new Job);
    j->id = S->next_job++;
    j->scene = scene_id;
    j->camera = camera ? *camera : S->scenes[scene_id]->desc.camera;
    j->params = *p;
    j->info.state = DL_JOB_QUEUED;
    j->info.lane = -1;
    j->info.t_submit = S->now();
    *job_id = j->id;
    S->queue.push_back(j->id);
    S->jobs[j->id] = std::move(j);
    S->cv.notify_all();
    return DL_OK;
}

dl_status dl_hip_service_poll(void* h, uint64_t job_id, dl_job_info* info) {
    Service* S = (Service*)h;
    if (!S || !info) return err(S, DL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(S->m);
    auto it = S->jobs.find(job_id);
    if (it == S->jobs.end()) return err(S, DL_ERR_INVALID, "unknown job id " + std::to_string(job_id));
    *info = it->second->info;
    if (info->state == DL_JOB_FAILED) S->err = it->second->error;
    return DL_OK;
}

dl_status dl_hip_service_wait(void* h, uint64_t job_id, dl_job_info* info) {
    Service* S = (Service*)h;
    if (!S) return err(S, DL_ERR_INVALID, "null argument");
    std::unique_lock<std::mutex> lk(S->m);
    auto it = S->jobs.find(job_id);
    if (it == S->jobs.end()) return err(S, DL_ERR_INVALID, "unknown job id " + std::to_string(job_id));
    Job* j = it->second.get();
    S->cv.wait(lk, [&] { return j->info.state >= DL_JOB_DONE; });
    if (info) *info = j->info;
    if (j->info.state == DL_JOB_FAILED) S->err = j->error;
    return DL_OK;
}

dl_status dl_hip_service_result(void* h, uint64_t job_id, float* rgb) {
    Service* S = (Service*)h;
    if (!S || !rgb) return err(S, DL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(S->m);
    auto it = S->jobs.find(job_id);
    if (it == S->jobs.end()) return err(S, DL_ERR_INVALID, "unknown job id " + std::to_string(job_id));
    Job& j = *it->second;
    if (j.info.state != DL_JOB_DONE)
        return err(S, DL_ERR_STATE, j.info.state == DL_JOB_FAILED ? "job failed: " + j.error : "job is not done");
    if (j.rgb.empty()) return err(S, DL_ERR_STATE, "the result was already taken");
    std::memcpy(rgb, j.rgb.data(), j.rgb.size() * sizeof(float));
    std::vector<float>().swap(j.rgb);
    return DL_OK;
}

dl_status dl_hip_service_cancel(void* h, uint64_t job_id) {
    Service* S = (Service*)h;
    if (!S) return err(S, DL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(S->m);
    auto it = S->jobs.find(job_id);
    if (it == S->jobs.end()) return err(S, DL_ERR_INVALID, "unknown job id " + std::to_string(job_id));
    Job& j = *it->second;
    if (j.info.state == DL_JOB_QUEUED) {
        S->queue.erase(std::find(S->queue.begin(), S->queue.end(), job_id));
        finish(*S, j, DL_JOB_CANCELLED, "");
        S->cv.notify_all();
    } else if (j.info.state == DL_JOB_RUNNING) {
        j.cancel = true;  // read by the lane before its next pass
    }
    return DL_OK;
}

dl_status dl_hip_service_stats(void* h, dl_service_stats* st) {
    Service* S = (Service*)h;
    if (!S || !st) return err(S, DL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(S->m);
    *st = S->stats;
    return DL_OK;
}

dl_status dl_hip_service_stop(void* h) {
    Service* S = (Service*)h;
    if (!S) return DL_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(S->m);
        S->stopping = true;
        for (uint64_t id : S->queue) finish(*S, *S->jobs[id], DL_JOB_CANCELLED, "");
        for (auto& kv : S->jobs)
            if (kv.second->info.state == DL_JOB_RUNNING) kv.second->cancel = true;
        S->queue.clear();
        S->cv.notify_all();
    }
    for (auto& l : S->lanes)
        if (l->th.joinable()) l->th.join();
    for (auto& l : S->lanes) dl_hip_destroy(l->ctx);
    delete S;
    return DL_OK;
}

const char* dl_hip_service_last_error(void* h) {
    Service* S = (Service*)h;
    return S ? S->err.c_str() : g_service_error.c_str();
}

}  // extern "C"

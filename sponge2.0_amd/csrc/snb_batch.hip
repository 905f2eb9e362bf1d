// snb_batch.hip -- one dispatch per kernel for the same step of several
// contexts.
//
// A step batch records the kernel launches that the ordinary entry points
// make for each member context (snb_batch_record_begin / _end around the
// calls), then merges them: launches of a kernel that has a batched form go
// out as one dispatch over an argument table with one entry per member
// (SnbSegTab), the rest as the single-context path makes them.  Each
// member's launches keep their order (the reciprocal pipeline's on a second
// stream, ordered against the rest as a context's own second stream is), so
// a member's arithmetic is that of its own launches: its results are
// bit-identical to running it alone.
#include <algorithm>
#include <map>

#include "snb_internal.h"

thread_local SnbRecorder* snb_rec = nullptr;

namespace
{
struct PlanOp
{
    std::function<void(hipStream_t)> solo;  // set for a single launch
    const void* bfn = nullptr;              // batched kernel
    unsigned grid = 0, block = 0;
    size_t lds = 0;
    size_t args_off = 0, blk_off = 0;  // offsets into the table buffer
    int nseg = 0, stride = 0;
    int lane = 0;
    bool join = false, writes_inputs = false;
};

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }
}  // namespace

struct snb_batch
{
    int device = 0;
    std::vector<snb_context*> member;
    std::vector<SnbRecorder> rec;
    int recording = -1;  // member being recorded
    bool planned = false;
    std::vector<PlanOp> plan;
    char* dtab = nullptr;
    size_t dtab_bytes = 0;
    int launches = 0, batched = 0, merged = 0;
    // the reciprocal lane's stream and the fork / join events
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
};

int snb_batch_create(snb_batch** out, int device)
{
    if (!out) return SNB_ERR_ARGUMENT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
        return SNB_ERR_ARGUMENT;
    snb_batch* b = new snb_batch();
    b->device = device;
    DeviceGuard g(device);
    if (hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming) != hipSuccess)
    {
        snb_batch_destroy(b);
        return SNB_ERR_HIP;
    }
    *out = b;
    return SNB_OK;
}

int snb_batch_destroy(snb_batch* b)
{
    if (!b) return SNB_OK;
    if (b->recording >= 0) snb_rec = nullptr;
    int rc = SNB_OK;
    DeviceGuard g(b->device);
    if (b->dtab && hipFree(b->dtab) != hipSuccess) rc = SNB_ERR_HIP;
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    if (b->ev_join) (void)hipEventDestroy(b->ev_join);
    if (b->side) (void)hipStreamDestroy(b->side);
    delete b;
    return rc;
}

int snb_batch_record_begin(snb_batch* b, snb_context* c)
{
    if (!b || !c) return SNB_ERR_ARGUMENT;
    if (b->recording >= 0 || b->planned || snb_rec) return SNB_ERR_ARGUMENT;
    if (c->device != b->device) return SNB_ERR_ARGUMENT;
    // one stream per member: the reciprocal pipeline on a second stream, and
    // the profiling events, are stream operations that are not launches
    if (c->overlap || c->prof.on) return SNB_ERR_UNSUPPORTED;
    for (snb_context* m : b->member)
        if (m == c) return SNB_ERR_ARGUMENT;
    b->member.push_back(c);
    b->rec.emplace_back();
    b->rec.back().weight = (double)c->n;
    b->recording = (int)b->member.size() - 1;
    snb_rec = &b->rec.back();
    return SNB_OK;
}

int snb_batch_record_end(snb_batch* b)
{
    if (!b || b->recording < 0) return SNB_ERR_ARGUMENT;
    snb_rec = nullptr;
    b->recording = -1;
    return SNB_OK;
}

static int batch_plan(snb_batch* b);

int snb_batch_plan(snb_batch* b, int* launches, int* batched)
{
    if (!b || b->recording >= 0 || b->planned) return SNB_ERR_ARGUMENT;
    b->plan.clear();
    b->launches = b->batched = b->merged = 0;
    const int rc = batch_plan(b);
    if (rc)
    {
        // nothing of a failed plan is kept; the records stay for the caller
        // to destroy the batch
        b->plan.clear();
        b->launches = b->batched = b->merged = 0;
        return rc;
    }
    if (launches) *launches = b->launches;
    if (batched) *batched = b->batched;
    return SNB_OK;
}

static int batch_plan(snb_batch* b)
{
    const int M = (int)b->member.size();
    std::vector<size_t> cur(M, 0);
    std::vector<char> host;
    auto place = [&](const void* p, size_t bytes, size_t align) {
        size_t off = align_up(host.size(), align);
        host.resize(off);
        host.insert(host.end(), (const char*)p, (const char*)p + bytes);
        return off;
    };
    std::map<const void*, size_t> lds_need;
    for (;;)
    {
        size_t issued = 0;  // every pass advances at least one member
        for (int m = 0; m < M; m++) issued += cur[m];
        // launches without a batched form go out as they are, in each
        // member's order
        for (int m = 0; m < M; m++)
        {
            std::vector<SnbRecOp>& ops = b->rec[m].ops;
            while (cur[m] < ops.size() && !ops[cur[m]].bfn)
            {
                PlanOp p;
                p.solo = ops[cur[m]].solo;
                p.lane = ops[cur[m]].lane;
                p.join = ops[cur[m]].join;
                p.writes_inputs = ops[cur[m]].writes_inputs;
                b->plan.push_back(std::move(p));
                b->launches++;
                cur[m]++;
            }
        }
        // Next dispatch: among the heads of the members that have issued the
        // fewest launches so far, the batched kernel (with its block size and
        // lane) at the head of most members.  Members whose sequences differ for a while
        // (plain and windowed pair path, grids with and without the fused
        // plane transform) so fall back into step with the others.
        size_t lag = (size_t)-1;
        for (int m = 0; m < M; m++)
            if (cur[m] < b->rec[m].ops.size()) lag = std::min(lag, cur[m]);
        if (lag == (size_t)-1) break;
        std::vector<std::tuple<const void*, unsigned, int>> keys;
        std::vector<int> votes;
        for (int m = 0; m < M; m++)
        {
            if (cur[m] >= b->rec[m].ops.size()) continue;
            const SnbRecOp& o = b->rec[m].ops[cur[m]];
            auto key = std::make_tuple(o.bfn, o.block, o.lane);
            auto it = std::find(keys.begin(), keys.end(), key);
            if (it == keys.end())
            {
                keys.push_back(key);
                votes.push_back(0);
                it = keys.end() - 1;
            }
            votes[it - keys.begin()] += 1;
        }
        int best = -1;
        for (int m = 0; m < M; m++)
        {
            if (cur[m] != lag || cur[m] >= b->rec[m].ops.size()) continue;
            const SnbRecOp& o = b->rec[m].ops[cur[m]];
            const int k = (int)(std::find(keys.begin(), keys.end(),
                                          std::make_tuple(o.bfn, o.block, o.lane)) -
                                keys.begin());
            if (best < 0 || votes[k] > votes[best]) best = k;
        }
        const auto key = keys[best];
        PlanOp p;
        p.bfn = std::get<0>(key);
        p.block = std::get<1>(key);
        p.lane = std::get<2>(key);
        std::vector<int> blk0(1, 0);
        std::vector<char> args;
        // the members' workgroups, heaviest first: the dispatch's tail is
        // then made of light workgroups
        std::vector<int> take;
        for (int m = 0; m < M; m++)
        {
            if (cur[m] >= b->rec[m].ops.size()) continue;
            const SnbRecOp& o = b->rec[m].ops[cur[m]];
            if (std::make_tuple(o.bfn, o.block, o.lane) != key) continue;
            take.push_back(m);
        }
        std::stable_sort(take.begin(), take.end(), [&](int x, int y) {
            return b->rec[x].ops[cur[x]].weight > b->rec[y].ops[cur[y]].weight;
        });
        for (int m : take)
        {
            const SnbRecOp& o = b->rec[m].ops[cur[m]];
            cur[m]++;
            if (o.grid == 0) continue;
            if (p.nseg == 0)
                p.stride = (int)o.args.size();
            else if ((int)o.args.size() != p.stride)
                return SNB_ERR_ARGUMENT;  // one kernel, one argument block
            p.join = p.join || o.join;
            p.writes_inputs = p.writes_inputs || o.writes_inputs;
            args.insert(args.end(), o.args.begin(), o.args.end());
            blk0.push_back(blk0.back() + (int)o.grid);
            p.lds = std::max(p.lds, o.lds);
            p.nseg++;
        }
        size_t now = 0;
        for (int m = 0; m < M; m++) now += cur[m];
        if (now == issued) return SNB_ERR_ARGUMENT;
        if (p.nseg == 0) continue;
        p.grid = (unsigned)blk0.back();
        p.args_off = place(args.data(), args.size(), 256);
        p.blk_off = place(blk0.data(), sizeof(int) * blk0.size(), 16);
        lds_need[p.bfn] = std::max(lds_need[p.bfn], p.lds);
        b->launches++;
        b->batched++;
        b->merged += p.nseg;
        b->plan.push_back(std::move(p));
    }
    DeviceGuard g(b->device);
    int rc = SNB_OK;
    // dynamic LDS above 64 KiB is granted per kernel: the largest request
    // of any plan so far (plans are made from one thread at a time)
    static std::map<std::pair<int, const void*>, size_t> granted;
    for (const auto& kv : lds_need)
    {
        size_t& g = granted[std::make_pair(b->device, kv.first)];
        if (kv.second > 65536 && kv.second > g)
        {
            if (hipFuncSetAttribute(kv.first, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)kv.second) != hipSuccess)
                rc = SNB_ERR_HIP;
            else
                g = kv.second;
        }
    }
    if (!rc && !host.empty())
    {
        if (hipMalloc((void**)&b->dtab, host.size()) != hipSuccess ||
            hipMemcpy(b->dtab, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess)
            rc = SNB_ERR_HIP;
        else
            b->dtab_bytes = host.size();
    }
    if (rc) return rc;
    b->planned = true;
    b->rec.clear();  // the plan holds what it needs
    return SNB_OK;
}

int snb_batch_launch(snb_batch* b, void* hip_stream)
{
    if (!b || !b->planned) return SNB_ERR_ARGUMENT;
    hipStream_t s = (hipStream_t)hip_stream;
    DeviceGuard g(b->device);
    // Lane 1 runs on the side stream.  ev_fork is recorded on the main stream
    // at the start and right after every launch that writes coordinates; the
    // side stream waits for its latest record before the next lane-1 launch.
    // A lane-0 launch marked join waits for everything on the side stream.
    SNB_CHECK(hipEventRecord(b->ev_fork, s));
    bool wait_due = true, side_open = false;
    for (const PlanOp& p : b->plan)
    {
        hipStream_t q = s;
        if (p.lane == 1)
        {
            if (wait_due)
            {
                SNB_CHECK(hipStreamWaitEvent(b->side, b->ev_fork, 0));
                wait_due = false;
            }
            side_open = true;
            q = b->side;
        }
        else if (p.join && side_open)
        {
            SNB_CHECK(hipEventRecord(b->ev_join, b->side));
            SNB_CHECK(hipStreamWaitEvent(s, b->ev_join, 0));
            side_open = false;
        }
        if (p.solo)
            p.solo(q);
        else
        {
            SnbSegTab t;
            t.args = b->dtab + p.args_off;
            t.blk0 = (const int*)(b->dtab + p.blk_off);
            t.nseg = p.nseg;
            t.stride = p.stride;
            void* kargs[] = {&t};
            SNB_CHECK(hipLaunchKernel(p.bfn, dim3(p.grid), dim3(p.block), kargs, p.lds, q));
        }
        if (p.lane == 0 && p.writes_inputs)
        {
            SNB_CHECK(hipEventRecord(b->ev_fork, s));
            wait_due = true;
        }
    }
    // the side stream's last launches rejoin the main stream (a captured
    // graph needs every forked stream joined)
    if (side_open)
    {
        SNB_CHECK(hipEventRecord(b->ev_join, b->side));
        SNB_CHECK(hipStreamWaitEvent(s, b->ev_join, 0));
    }
    SNB_CHECK(hipGetLastError());
    return SNB_OK;
}

int snb_batch_info(snb_batch* b, int* members, int* launches, int* batched, int* merged)
{
    if (!b) return SNB_ERR_ARGUMENT;
    if (members) *members = (int)b->member.size();
    if (launches) *launches = b->launches;
    if (batched) *batched = b->batched;
    if (merged) *merged = b->merged;
    return SNB_OK;
}

// Scheduler and runtime of the CPU HIP emulation (test harness; see hip/hip_runtime.h).
// Lanes are ucontext coroutines. A lane runs until it reaches a wave collective or exits;
// when every lane of the resident blocks waits, each wave resolves the group of lanes waiting
// at its earliest site (sites in headers before sites in .hip files, then the lower line),
// which is the order in which structured SIMT control flow reaches them: the lanes inside a
// branch meet a collective there before the others meet the one after the branch. A
// collective's result covers only the lanes of its group, as inactive lanes on the GPU.
// Kernel arguments that are scene records are checked before each launch (hybrid cycle
// fields and mix-in slots in range), a watchdog aborts a launch that runs longer than
// EMU_WATCHDOG seconds (default 120), and a launch in which no lane can move aborts.
#include <hip/hip_runtime.h>
#include <signal.h>
#include <sys/mman.h>
#include <ucontext.h>
#include <unistd.h>

#include <chrono>
#include <map>

#include "nd_dev.h"
#include "nd_hip.h"

EmuIdx* emu_cur = nullptr;
dim3 emu_grid_dim, emu_block_dim;

namespace {
struct Lane : EmuIdx {
    ucontext_t ctx;
    char* stack = nullptr;
    bool done = false, blocked = false;
    int kind = 0, arg = 0, line = 0;
    const char* file = "";
    uint64_t val = 0, res = 0;
};
struct Block { std::vector<Lane> lanes; };
constexpr size_t kStack = 512 * 1024;
ucontext_t g_sched;
Lane* g_lane = nullptr;
std::function<void()>* g_body = nullptr;
const char* g_kernel = "";
std::map<char*, std::pair<char*, size_t>> g_allocs;
uint64_t g_clock = 0;
uint64_t g_swaps = 0;

void trampoline() {
    (*g_body)();
    g_lane->done = true;
    swapcontext(&g_lane->ctx, &g_sched);
}
void on_alarm(int) {
    // async-signal-safe: write(2) of fixed text and the kernel name, then abort
    static const char msg[] = "\n[emu] WATCHDOG: a launch ran past EMU_WATCHDOG seconds: ";
    (void)!write(2, msg, sizeof(msg) - 1);
    (void)!write(2, g_kernel, strlen(g_kernel));
    (void)!write(2, "\n", 1);
    abort();
}
int file_rank(const char* f) {
    size_t n = strlen(f);
    return (n > 4 && strcmp(f + n - 4, ".hpp") == 0) ? 0 : 1;
}
bool site_less(const Lane* a, const Lane* b) {
    int ra = file_rank(a->file), rb = file_rank(b->file);
    if (ra != rb) return ra < rb;
    int c = strcmp(a->file, b->file);
    if (c) return c < 0;
    return a->line < b->line;
}
bool same_site(const Lane* a, const Lane* b) { return a->line == b->line && strcmp(a->file, b->file) == 0 && a->kind == b->kind; }
}  // namespace

uint64_t emu_clock() { return ++g_clock; }

uint64_t emu_collective(int kind, uint64_t v, int arg, const char* file, int line) {
    Lane* L = g_lane;
    L->kind = kind; L->val = v; L->arg = arg; L->file = file; L->line = line;
    L->blocked = true;
    ++g_swaps;
    swapcontext(&L->ctx, &g_sched);
    emu_cur = L;
    return L->res;
}

void emu_run(dim3 grid, dim3 block, const char* name, std::function<void()> body) {
    static bool sig = false;
    if (!sig) { signal(SIGALRM, on_alarm); sig = true; }
    const char* wd = getenv("EMU_WATCHDOG");
    alarm(wd ? atoi(wd) : 120);
    g_kernel = name;
    emu_grid_dim = grid;
    emu_block_dim = block;
    g_body = &body;
    const uint32_t nthreads = block.x * block.y * block.z;
    const uint64_t nblocks = (uint64_t)grid.x * grid.y * grid.z;
    const char* rs = getenv("EMU_RESIDENT");
    const size_t R = rs ? (size_t)atoi(rs) : 16;
    std::vector<Block*> act;
    uint64_t next = 0;
    const bool verbose = getenv("EMU_VERBOSE") != nullptr;
    if (verbose) fprintf(stderr, "[emu] launch %s grid %u block %u\n", name, (unsigned)nblocks, nthreads);
    while (next < nblocks || !act.empty()) {
        while (act.size() < R && next < nblocks) {
            Block* B = new Block();
            B->lanes.resize(nthreads);
            const uint64_t b = next++;
            for (uint32_t i = 0; i < nthreads; ++i) {
                Lane& L = B->lanes[i];
                L.tid = dim3(i % block.x, (i / block.x) % block.y, i / (block.x * block.y));
                L.bid = dim3((uint32_t)(b % grid.x), (uint32_t)((b / grid.x) % grid.y), (uint32_t)(b / ((uint64_t)grid.x * grid.y)));
                L.lane = (int)(i % 64);
                L.stack = (char*)mmap(nullptr, kStack, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
                getcontext(&L.ctx);
                L.ctx.uc_stack.ss_sp = L.stack;
                L.ctx.uc_stack.ss_size = kStack;
                L.ctx.uc_link = nullptr;
                makecontext(&L.ctx, trampoline, 0);
            }
            act.push_back(B);
        }
        bool ran = false;
        for (Block* B : act)
            for (Lane& L : B->lanes)
                if (!L.done && !L.blocked) {
                    g_lane = &L;
                    emu_cur = &L;
                    swapcontext(&g_sched, &L.ctx);
                    ran = true;
                }
        bool resolved = false;
        for (Block* B : act) {
            const uint32_t nw = (nthreads + 63) / 64;
            for (uint32_t w = 0; w < nw; ++w) {
                Lane* first = nullptr;
                for (uint32_t i = w * 64; i < std::min(nthreads, (w + 1) * 64); ++i) {
                    Lane& L = B->lanes[i];
                    if (L.done || !L.blocked || L.kind == EMU_SYNC) continue;
                    if (!first || site_less(&L, first)) first = &L;
                }
                if (!first) continue;
                std::vector<Lane*> g;
                for (uint32_t i = w * 64; i < std::min(nthreads, (w + 1) * 64); ++i) {
                    Lane& L = B->lanes[i];
                    if (!L.done && L.blocked && L.kind != EMU_SYNC && same_site(&L, first)) g.push_back(&L);
                }
                uint64_t ballot = 0;
                for (Lane* L : g) ballot |= (L->val & 1u) << L->lane;
                for (Lane* L : g) {
                    if (first->kind == EMU_BALLOT) L->res = ballot;
                    else if (first->kind == EMU_FIRST) L->res = g[0]->val;
                    else {  // shuffles: down, up, by index
                        int src = first->kind == EMU_SHFL ? L->lane + L->arg : first->kind == EMU_SHFL_UP ? L->lane - L->arg : L->arg;
                        L->res = L->val;
                        for (Lane* S : g) if (S->lane == src) L->res = S->val;
                    }
                    L->blocked = false;
                }
                resolved = true;
            }
            bool all_sync = true, any = false;
            for (Lane& L : B->lanes) if (!L.done) { any = true; if (!(L.blocked && L.kind == EMU_SYNC)) all_sync = false; }
            if (any && all_sync) { for (Lane& L : B->lanes) if (!L.done) L.blocked = false; resolved = true; }
        }
        for (size_t k = 0; k < act.size();) {
            bool all = true;
            for (Lane& L : act[k]->lanes) all = all && L.done;
            if (all) {
                for (Lane& L : act[k]->lanes) munmap(L.stack, kStack);
                delete act[k];
                act.erase(act.begin() + k);
            } else {
                ++k;
            }
        }
        if (!ran && !resolved && !act.empty()) {
            fprintf(stderr, "[emu] DEADLOCK in %s\n", name);
            abort();
        }
    }
    alarm(0);
    g_lane = nullptr;
}

hipError_t hipMalloc(void** p, size_t n) {
    const size_t pg = 4096;
    const size_t body = ((n ? n : 1) + pg - 1) / pg * pg;
    char* base = (char*)mmap(nullptr, body + pg, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
    if (base == MAP_FAILED) return hipErrorOutOfMemory;
    mprotect(base + body, pg, PROT_NONE);
    const size_t n16 = (n + 15) / 16 * 16;
    char* ptr = base + body - n16;
    g_allocs[ptr] = {base, body + pg};
    *p = ptr;
    return hipSuccess;
}
hipError_t hipFree(void* p) {
    if (!p) return hipSuccess;
    auto it = g_allocs.find((char*)p);
    if (it == g_allocs.end()) { fprintf(stderr, "[emu] hipFree of unknown %p\n", p); abort(); }
    munmap(it->second.first, it->second.second);
    g_allocs.erase(it);
    return hipSuccess;
}
hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int) {
    const char* cu = getenv("EMU_CUS");
    p->multiProcessorCount = cu ? atoi(cu) : 8;
    p->totalGlobalMem = (size_t)8 << 30;
    snprintf(p->gcnArchName, sizeof(p->gcnArchName), "gfx950-emu");
    snprintf(p->name, sizeof(p->name), "emu");
    return hipSuccess;
}
hipError_t hipGetSymbolAddress(void** p, const void* sym) { *p = (void*)sym; return hipSuccess; }
namespace { struct Ev { std::chrono::steady_clock::time_point t; }; }
hipError_t hipEventCreate(hipEvent_t* e) { *e = new Ev(); return hipSuccess; }
hipError_t hipEventDestroy(hipEvent_t e) { delete (Ev*)e; return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) { ((Ev*)e)->t = std::chrono::steady_clock::now(); return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
    *ms = std::chrono::duration<float, std::milli>(((Ev*)b)->t - ((Ev*)a)->t).count();
    return hipSuccess;
}

void emu_check_scene(const DevScene* s) {
    if (!s) return;
    if (g_allocs.empty()) return;
    if (s->count < 1 || s->count > ND_DEV_MAX_INST) { fprintf(stderr, "[emu] scene count %d\n", s->count); abort(); }
    for (int i = 0; i < s->count; ++i) {
        const DevInst& in = s->inst[i];
        if (!in.hybrid) continue;
        bool bad = in.mix_count < 0 || in.mix_count > ND_MAX_MIXINS || in.hyb_cycle < 1 || in.hyb_total < 0 || in.hyb_base_n < 0;
        int sum = in.hyb_base_n;
        for (int j = 0; j < in.mix_count && !bad; ++j) {
            const DevMixin& m = in.mix[1 + j];
            if (m.iters < 0 || m.formula < 0 || m.formula >= ND_FORMULA_COUNT) bad = true;
            sum += m.iters;
        }
        if (in.mix[0].formula < 0 || in.mix[0].formula >= ND_FORMULA_COUNT) bad = true;
        if (getenv("EMU_VERBOSE") || bad)
            fprintf(stderr, "[emu] inst %d hybrid: mix_count %d base_n %d total %d cycle %d (sum %d) formulas %d/%d/%d/%d iters %d/%d/%d\n",
                    i, in.mix_count, in.hyb_base_n, in.hyb_total, in.hyb_cycle, sum, in.mix[0].formula, in.mix[1].formula,
                    in.mix[2].formula, in.mix[3].formula, in.mix[1].iters, in.mix[2].iters, in.mix[3].iters);
        if (bad) abort();
    }
}
void emu_check_job(const DevJob* j) { if (j && !g_allocs.empty()) emu_check_scene(&j->scene); }

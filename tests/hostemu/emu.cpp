// TEST INFRASTRUCTURE: fiber scheduler and guard-paged allocator of the host
// emulation (see hip/hip_runtime.h).
#include "hip/hip_runtime.h"
namespace emu {
dim3 g_block, g_grid, g_blockIdx;
int g_cur = 0;
bool g_in_fiber = false;
ucontext_t g_sched;
std::vector<Fiber> g_fibers;
double g_shfl[1024];
std::function<void()> g_body;
static const size_t kStack = 256 * 1024;
static void entry() {
  g_body();
  g_fibers[g_cur].done = true;
  swapcontext(&g_fibers[g_cur].ctx, &g_sched);
}
void launch(dim3 grid, dim3 block, std::function<void()> body) {
  g_grid = grid;
  g_block = block;
  g_body = body;
  const int nt = (int)(block.x * block.y * block.z);
  if ((int)g_fibers.size() < nt) {
    size_t old = g_fibers.size();
    g_fibers.resize(nt);
    for (size_t i = old; i < (size_t)nt; i++) g_fibers[i].stack = (char*)malloc(kStack);
  }
  for (unsigned bz = 0; bz < grid.z; bz++)
    for (unsigned by = 0; by < grid.y; by++)
      for (unsigned bx = 0; bx < grid.x; bx++) {
        g_blockIdx = dim3(bx, by, bz);
        for (int t = 0; t < nt; t++) {
          Fiber& f = g_fibers[t];
          f.done = false;
          getcontext(&f.ctx);
          f.ctx.uc_stack.ss_sp = f.stack;
          f.ctx.uc_stack.ss_size = kStack;
          f.ctx.uc_link = nullptr;
          makecontext(&f.ctx, entry, 0);
        }
        g_in_fiber = true;
        bool left = true;
        while (left) {
          left = false;
          for (int t = 0; t < nt; t++) {
            if (g_fibers[t].done) continue;
            g_cur = t;
            swapcontext(&g_sched, &g_fibers[t].ctx);
            if (!g_fibers[t].done) left = true;
          }
        }
        g_in_fiber = false;
      }
}
}  // namespace emu

struct Map {
  char* base;
  size_t len;
  size_t n;  // bytes asked for
};
static std::map<void*, Map> g_maps;
hipError_t hipMalloc(void** p, size_t n) {
  const size_t pg = (size_t)sysconf(_SC_PAGESIZE);
  const size_t data = (n + pg - 1) / pg * pg;
  const size_t len = data + 2 * pg;
  char* base = (char*)mmap(nullptr, len, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
  if (base == MAP_FAILED) return 2;
  mprotect(base, pg, PROT_NONE);
  mprotect(base + pg + data, pg, PROT_NONE);
  // EMU_ALIGN=start: buffer begins right after the front guard page;
  // otherwise it ends right before the back guard page
  const char* al = getenv("EMU_ALIGN");
  char* q = (al && !strcmp(al, "start")) ? base + pg : base + pg + data - ((n + 7) / 8 * 8);
  memset(q, 0xff, n);  // NaN garbage like fresh device memory might hold
  g_maps[q] = Map{base, len, n};
  *p = q;
  return hipSuccess;
}
hipError_t hipFree(void* p) {
  if (!p) return hipSuccess;
  auto it = g_maps.find(p);
  if (it == g_maps.end()) return 1;
  munmap(it->second.base, it->second.len);
  g_maps.erase(it);
  return hipSuccess;
}
hipError_t hipMemGetAddressRange(hipDeviceptr_t* base, size_t* size, hipDeviceptr_t p) {
  const char* c = (const char*)p;
  auto it = g_maps.upper_bound(p);
  if (it != g_maps.begin()) {
    --it;
    const char* q = (const char*)it->first;
    if (c >= q && c < q + it->second.n) {
      *base = it->first;
      *size = it->second.n;
      return hipSuccess;
    }
  }
  *base = p;
  *size = (size_t)1 << 62;
  return hipSuccess;
}

// The static device scratch of kc_rank_list_noalloc.
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops) rank lists through kc_rank_list_noalloc, whose scratch
// comes from this file-scope array instead of the context's arena.  Every GPU
// the library runs on holds its own copy until the process ends; host code
// never uses the array's name (on the host it is only a stand-in symbol), but
// the address hipGetSymbolAddress returns for the current device, looked up
// once per device and kept here.
#include <mutex>

#include "common.hpp"

namespace kc {

namespace {

// Enough for the device-sized ranking of any list or forest of at most
// KC_NOALLOC_MAX_N elements with the built-in tuning (at most 176 bytes per
// element, reached by a forest of singletons); see kc_rank_list_noalloc.
constexpr size_t kNoAllocScratchBytes = size_t{192} << 20;
static_assert(kNoAllocScratchBytes == KC_NOALLOC_SCRATCH_BYTES, "header and array disagree");

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
char* g_ptr[kMaxDevices] = {};
bool g_looked_up[kMaxDevices] = {};

}  // namespace

bool noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return false;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_looked_up[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
    }
    g_ptr[device] = static_cast<char*>(p);
    g_looked_up[device] = true;
  }
  return g_ptr[device] != nullptr;
}

bool noalloc_scratch(int device, char** base, size_t* bytes) {
  if (device < 0 || device >= kMaxDevices) return false;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[device]) return false;
  *base = g_ptr[device];
  *bytes = kNoAllocScratchBytes;
  return true;
}

}  // namespace kc

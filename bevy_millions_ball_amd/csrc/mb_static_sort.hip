// mb_static_sort.hip -- mb_radix_sort_static: the stage entry point
// mb_radix_sort with the library's own static device workspace, for callers
// that cannot allocate device memory.
//
// The workspace is the file-scope `scratch` array below, one copy per GPU the
// library runs on, held until the process ends.  Host code never passes the
// array's name to a kernel or a copy: it takes the array's device address
// with hipGetSymbolAddress on the entry point's first call on each device
// (with that device current) and uses that address.  The sort kernels are the
// ones mb_radix_sort launches (mb_kernels.hip), with the same workspace
// layout for MB_RADIX_SORT_STATIC_MAX_KEYS keys.
#include <mutex>

#include "../../include/millions_ball.h"
#include "mb_internal.h"

namespace {

constexpr size_t align256(size_t x) { return (x + 255u) & ~(size_t)255u; }

// Bytes of the sort workspace for max_keys keys: the layout of
// mb::sort_workspace_layout (digit totals, per-(digit, tile) counts, the
// keys and values between passes), each part 256-byte aligned.
constexpr size_t static_sort_bytes(uint32_t max_keys)
{
    const size_t tiles = (max_keys + mb::kSortTile - 1u) / mb::kSortTile;
    const size_t counts = align256(mb::kRadix * 4u);
    const size_t keys = align256(counts + (size_t)mb::kRadix * tiles * 4u);
    const size_t vals = align256(keys + (size_t)max_keys * 4u);
    return align256(vals + (size_t)max_keys * 4u);
}

constexpr size_t kScratchBytes = static_sort_bytes(MB_RADIX_SORT_STATIC_MAX_KEYS);

// 8.25 MiB per GPU for 2^20 keys.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch_ptr[kMaxDevices] = {};

// The device address of the current device's `scratch`, looked up once per
// device; null (with the error set) on failure.
void *scratch_address(int device)
{
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[device]) return g_scratch_ptr[device];
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr) {
        mb::set_error(MB_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        return nullptr;
    }
    g_scratch_ptr[device] = ptr;
    return ptr;
}

}  // namespace

extern "C" int mb_radix_sort_static(const uint32_t *keys_in, const uint32_t *vals_in,
                                    uint32_t *keys_out, uint32_t *vals_out, uint32_t n,
                                    uint32_t pass_begin, uint32_t pass_end, void *stream)
{
    if (n == 0) return MB_OK;
    if (n > MB_RADIX_SORT_STATIC_MAX_KEYS)
        return mb::set_error(MB_ERR_INVALID_ARGUMENT,
                             "number_of_keys exceeds MB_RADIX_SORT_STATIC_MAX_KEYS (2^20); "
                             "use mb_radix_sort with a workspace");
    if (!keys_in || !keys_out || !vals_out)
        return mb::set_error(MB_ERR_INVALID_ARGUMENT, "null argument");
    if (pass_end > 4u || pass_begin > pass_end)
        return mb::set_error(MB_ERR_INVALID_ARGUMENT, "pass range must lie in 0..4");
    if (pass_begin == pass_end && vals_in == nullptr)
        return mb::set_error(MB_ERR_INVALID_ARGUMENT, "init_vals needs at least one pass");
    if (keys_in == keys_out || (vals_in && vals_in == vals_out))
        return mb::set_error(MB_ERR_INVALID_ARGUMENT, "input and output must not alias");
    int device = 0;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess) return mb::set_error(MB_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (device < 0 || device >= kMaxDevices) return mb::set_error(MB_ERR_HIP, "device index out of range");
    void *ws = scratch_address(device);
    if (!ws) return MB_ERR_HIP;
    const mb::SortWorkspaceLayout l = mb::sort_workspace_layout(MB_RADIX_SORT_STATIC_MAX_KEYS);
    if (l.bytes > kScratchBytes) return mb::set_error(MB_ERR_HIP, "static sort workspace too small");
    hipStream_t st = static_cast<hipStream_t>(stream);
    // The digit totals must start at zero; a previous call on this GPU left
    // them zero unless it stopped part way, so they are cleared by a kernel
    // (a memset node would misbehave in a captured graph).
    e = mb::launch_fill_u32(reinterpret_cast<uint32_t *>(static_cast<char *>(ws) + l.hist_offset), 0u,
                            mb::kRadix, st);
    if (e == hipSuccess)
        e = mb::launch_radix_sort(ws, MB_RADIX_SORT_STATIC_MAX_KEYS, keys_in, vals_in, keys_out,
                                  vals_out, n, pass_begin, pass_end, st);
    if (e != hipSuccess) return mb::set_error(MB_ERR_HIP, std::string("radix sort: ") + hipGetErrorString(e));
    return MB_OK;
}

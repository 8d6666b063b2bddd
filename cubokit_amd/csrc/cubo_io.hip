// cubo_io.hip — ingestion into HBM (include/cubo_hip.h, "ingestion"): the .npy header, files and host tiles
// copied to the device through pinned double buffers, and the streamed cube's ring.
//
// The reference reads a file through a read-only mapping and converts every element to Float32 on the CPU
// (Sources/CuboCore/MappedCube.swift:98-164), then copies tiles into a Metal buffer (ResidentCube.swift:49-55,
// TileFeed, MetalReducers.swift:494-573). Here the raw bytes go to the device: worker threads read chunk i + 1
// (pread from the file, or memcpy from host memory) into one pinned buffer while chunk i's DMA runs from the other;
// uint16 and Float32 land in place, every other type is converted to Float32 by a device kernel after its DMA.
#include <hip/hip_runtime.h>

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cubo_hip.h"
#include "cube.hpp"
#include "errors.hpp"
#include "host_math.hpp"
#include "reduce_kernels.hpp"

using cubo::DeviceGuard;
using cubo::element_size;
using cubo::fail;
using cubo::hip_fail;

namespace cubo {

constexpr size_t kFeedSlotBytes = (size_t)64 << 20;   // pinned bytes per slot

struct Feed {
    int device = 0;
    void *pinned[2] = {nullptr, nullptr};      // mapped: the device reads the slots over PCIe itself
    void *pinned_dev[2] = {nullptr, nullptr};  // their device addresses
    hipEvent_t copied[2] = {nullptr, nullptr};  // copy stream: the kernel that read the slot is done
    hipEvent_t consumed[2] = {nullptr, nullptr};
    hipStream_t copy = nullptr;
    int next = 0;
    // streamed cubes: the device ring
    void *ring[2] = {nullptr, nullptr};
    int64_t ring_row0[2] = {-1, -1}, ring_rows[2] = {0, 0};
    int64_t tile_rows = 0;
    int ring_next = 0;
};

namespace {

int element_bytes(int t) {
    switch (t) {
    case CUBO_HIP_ELEMENT_U8: case CUBO_HIP_ELEMENT_I8: return 1;
    case CUBO_HIP_ELEMENT_U16: case CUBO_HIP_ELEMENT_I16: return 2;
    case CUBO_HIP_ELEMENT_F32: case CUBO_HIP_ELEMENT_U32: case CUBO_HIP_ELEMENT_I32: return 4;
    case CUBO_HIP_ELEMENT_F64: case CUBO_HIP_ELEMENT_U64: case CUBO_HIP_ELEMENT_I64: return 8;
    default: return 0;
    }
}

// The cube type an element type is held as: uint16 stays, everything else is Float32.
int held_as(int t) { return t == CUBO_HIP_ELEMENT_U16 ? CUBO_HIP_U16 : CUBO_HIP_F32; }
bool native(int t, int dtype) {
    return (t == CUBO_HIP_ELEMENT_U16 && dtype == CUBO_HIP_U16) || (t == CUBO_HIP_ELEMENT_F32 && dtype == CUBO_HIP_F32);
}

// dst[i] = Float32(src[i]): the hardware conversions round to nearest even, as ElementType.convert
// (MappedCube.swift:56-83; vDSP and Swift's Float(_:)).
template <typename S>
__global__ void to_f32(const S *__restrict__ src, float *__restrict__ dst, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = (float)src[i];
}

hipError_t launch_to_f32(int t, const void *src, float *dst, int64_t n, hipStream_t s) {
    const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 8192);
    switch (t) {
#define CUBO_TO_F32(E, T)                                                                                \
    case E: hipLaunchKernelGGL(to_f32<T>, dim3(blocks), dim3(256), 0, s, (const T *)src, dst, n); break;
    CUBO_TO_F32(CUBO_HIP_ELEMENT_F64, double)
    CUBO_TO_F32(CUBO_HIP_ELEMENT_U8, uint8_t)
    CUBO_TO_F32(CUBO_HIP_ELEMENT_I8, int8_t)
    CUBO_TO_F32(CUBO_HIP_ELEMENT_U16, uint16_t)
    CUBO_TO_F32(CUBO_HIP_ELEMENT_I16, int16_t)
    CUBO_TO_F32(CUBO_HIP_ELEMENT_U32, uint32_t)
    CUBO_TO_F32(CUBO_HIP_ELEMENT_I32, int32_t)
    CUBO_TO_F32(CUBO_HIP_ELEMENT_U64, uint64_t)
    CUBO_TO_F32(CUBO_HIP_ELEMENT_I64, int64_t)
#undef CUBO_TO_F32
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int make_feed(cubo_hip_cube *c) {
    if (c->feed) return CUBO_HIP_OK;
    DeviceGuard g(c->device);
    Feed *f = new (std::nothrow) Feed();
    if (!f) return fail(CUBO_HIP_ERR_OUT_OF_MEMORY, "host allocation");
    f->device = c->device;
    hipError_t e = hipStreamCreateWithFlags(&f->copy, hipStreamNonBlocking);
    for (int s = 0; s < 2 && e == hipSuccess; ++s) {
        e = hipHostMalloc(&f->pinned[s], kFeedSlotBytes, hipHostMallocMapped);
        if (e == hipSuccess) e = hipHostGetDevicePointer(&f->pinned_dev[s], f->pinned[s], 0);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&f->copied[s], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&f->consumed[s], hipEventDisableTiming);
    }
    c->feed = f;
    if (e != hipSuccess) { free_feed(c); return hip_fail(e, "allocating the pinned feed"); }
    return CUBO_HIP_OK;
}

// Copies n bytes of raw elements of type t, produced by fill(pinned, src_offset, bytes), to dst (the cube's dtype)
// on the feed's copy stream, chunk by chunk through the two pinned slots. `after` (may be NULL) is an event the
// copies must wait for (the last reader of the destination). Returns once every chunk has been read from the source.
template <typename Fill>
int feed_bytes(cubo_hip_cube *c, void *dst, size_t nbytes, int t, hipEvent_t after, Fill fill) {
    Feed *f = c->feed;
    const int eb = element_bytes(t);
    const bool conv = !native(t, c->dtype);
    const int out_b = element_size(c->dtype);
    if (after) CUBO_HIP_TRY(hipStreamWaitEvent(f->copy, after, 0), "ordering the feed");
    const size_t chunk = (kFeedSlotBytes / (size_t)eb) * (size_t)eb;
    for (size_t off = 0; off < nbytes; off += chunk) {
        const size_t len = std::min(chunk, nbytes - off);
        const int s = f->next;
        f->next ^= 1;
        CUBO_HIP_TRY(hipEventSynchronize(f->copied[s]), "waiting for a pinned buffer");   // its last reader is done
        if (!fill(f->pinned[s], off, len)) return CUBO_HIP_ERR_FILE;                        // fill set the message
        char *out = (char *)dst + (off / eb) * out_b;
        // a kernel pulls the slot over PCIe (many reads in flight), converting on the way when the element type
        // is not the cube's
        if (!conv && ((uintptr_t)out & 15) == 0)   // the copy kernel moves 16-byte words
            CUBO_HIP_TRY(cubo::launch_copy_bytes(f->pinned_dev[s], out, len, f->copy), "copying to HBM");
        else if (!conv)
            CUBO_HIP_TRY(hipMemcpyAsync(out, f->pinned[s], len, hipMemcpyHostToDevice, f->copy), "copying to HBM");
        else
            CUBO_HIP_TRY(launch_to_f32(t, f->pinned_dev[s], (float *)out, (int64_t)(len / eb), f->copy), "converting");
        CUBO_HIP_TRY(hipEventRecord(f->copied[s], f->copy), "hipEventRecord");
    }
    // later work on the cube's stream waits for every copy of this call
    const int last = f->next ^ 1;
    CUBO_HIP_TRY(hipStreamWaitEvent(c->stream, f->copied[last], 0), "ordering the cube after the feed");
    return CUBO_HIP_OK;
}

// ---- the .npy header (NPYFile.swift:28-130) ----

struct Fields {
    bool has_descr = false, has_fortran = false, has_shape = false;
    std::string descr;
    bool fortran = false;
    std::vector<int64_t> shape;
};

int not_npy(const char *what) { return fail(CUBO_HIP_ERR_FILE, "notNPY: %s", what); }

// The header's Python dict literal, with the reference's grammar (NPYFile.swift:158-218).
int parse_dict(const std::string &text, Fields *out) {
    size_t b = 0, e = text.size();
    auto space = [](char ch) { return ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r'; };
    while (b < e && space(text[b])) ++b;
    while (e > b && space(text[e - 1])) --e;
    if (e - b < 2 || text[b] != '{' || text[e - 1] != '}') return not_npy("header is not a dict literal");
    std::string s = text.substr(b + 1, e - b - 2);
    size_t i = 0;
    auto skip = [&] { while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n')) ++i; };
    auto quoted = [&](std::string *v) -> int {
        skip();
        if (i >= s.size() || (s[i] != '\'' && s[i] != '"')) return not_npy("expected a quoted string");
        const char q = s[i++];
        const size_t end = s.find(q, i);
        if (end == std::string::npos) return not_npy("unterminated string");
        *v = s.substr(i, end - i);
        i = end + 1;
        return CUBO_HIP_OK;
    };
    auto expect = [&](char ch) -> int {
        skip();
        if (i >= s.size() || s[i] != ch) return fail(CUBO_HIP_ERR_FILE, "notNPY: expected '%c'", ch);
        ++i;
        return CUBO_HIP_OK;
    };
    while (true) {
        skip();
        if (i >= s.size()) break;
        std::string key;
        if (int st = quoted(&key)) return st;
        if (int st = expect(':')) return st;
        skip();
        if (key == "descr") {
            if (out->has_descr) return not_npy("'descr' twice");
            if (int st = quoted(&out->descr)) return st;
            out->has_descr = true;
        } else if (key == "fortran_order") {
            if (out->has_fortran) return not_npy("'fortran_order' twice");
            if (s.compare(i, 4, "True") == 0) { out->fortran = true; i += 4; }
            else if (s.compare(i, 5, "False") == 0) { out->fortran = false; i += 5; }
            else return not_npy("'fortran_order' is not True or False");
            out->has_fortran = true;
        } else if (key == "shape") {
            if (out->has_shape) return not_npy("'shape' twice");
            if (int st = expect('(')) return st;
            const size_t close = s.find(')', i);
            if (close == std::string::npos) return not_npy("unterminated 'shape'");
            const std::string inner = s.substr(i, close - i);
            i = close + 1;
            size_t p = 0;
            while (p <= inner.size()) {
                size_t q = inner.find(',', p);
                if (q == std::string::npos) q = inner.size();
                std::string part = inner.substr(p, q - p);
                size_t a = 0, z = part.size();
                while (a < z && (part[a] == ' ' || part[a] == '\t')) ++a;
                while (z > a && (part[z - 1] == ' ' || part[z - 1] == '\t')) --z;
                part = part.substr(a, z - a);
                if (!part.empty()) {
                    if (part.back() == 'L') part.pop_back();
                    errno = 0;
                    char *endp = nullptr;
                    const long long v = strtoll(part.c_str(), &endp, 10);
                    if (part.empty() || *endp != '\0' || errno == ERANGE)
                        return fail(CUBO_HIP_ERR_FILE, "notNPY: 'shape' holds %s", part.c_str());
                    out->shape.push_back((int64_t)v);
                }
                p = q + 1;
            }
            out->has_shape = true;
        } else {
            return fail(CUBO_HIP_ERR_FILE, "notNPY: unknown key '%s'", key.c_str());
        }
        skip();
        if (i < s.size() && s[i] == ',') { ++i; continue; }
        skip();
        if (i < s.size()) return fail(CUBO_HIP_ERR_FILE, "notNPY: expected ',' after '%s'", key.c_str());
    }
    return CUBO_HIP_OK;
}

// NPYFile.elementType (NPYFile.swift:73-97): little-endian or byte-order free; big-endian refused by name.
int element_type(const std::string &descr, int *t) {
    if (descr.size() < 3) return fail(CUBO_HIP_ERR_FILE, "unsupportedElementType: %s", descr.c_str());
    const char order = descr[0];
    const std::string code = descr.substr(1);
    if (order == '>') {
        if (code == "u1" || code == "i1") return element_type("|" + code, t);
        return fail(CUBO_HIP_ERR_FILE, "bigEndianNotSupported: %s", descr.c_str());
    }
    if (order != '<' && order != '|' && order != '=')
        return fail(CUBO_HIP_ERR_FILE, "unsupportedElementType: %s", descr.c_str());
    static const struct { const char *code; int t; } table[] = {
        {"f4", CUBO_HIP_ELEMENT_F32}, {"f8", CUBO_HIP_ELEMENT_F64}, {"u1", CUBO_HIP_ELEMENT_U8},
        {"i1", CUBO_HIP_ELEMENT_I8}, {"u2", CUBO_HIP_ELEMENT_U16}, {"i2", CUBO_HIP_ELEMENT_I16},
        {"u4", CUBO_HIP_ELEMENT_U32}, {"i4", CUBO_HIP_ELEMENT_I32}, {"u8", CUBO_HIP_ELEMENT_U64},
        {"i8", CUBO_HIP_ELEMENT_I64}};
    for (const auto &e : table)
        if (code == e.code) { *t = e.t; return CUBO_HIP_OK; }
    return fail(CUBO_HIP_ERR_FILE, "unsupportedElementType: %s", descr.c_str());
}

struct Fd {
    int fd = -1;
    ~Fd() { if (fd >= 0) close(fd); }
};

int open_file(const char *path, Fd *fd, int64_t *size) {
    fd->fd = open(path, O_RDONLY | O_CLOEXEC);
    if (fd->fd < 0) return fail(CUBO_HIP_ERR_FILE, "cannotOpen: %s: %s", path, strerror(errno));
    struct stat st;
    if (fstat(fd->fd, &st) != 0) return fail(CUBO_HIP_ERR_FILE, "cannotOpen: %s: %s", path, strerror(errno));
    *size = (int64_t)st.st_size;
    return CUBO_HIP_OK;
}

bool read_at(int fd, void *dst, size_t n, int64_t offset) {
    char *p = (char *)dst;
    while (n > 0) {
        const ssize_t r = pread(fd, p, n, (off_t)offset);
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) return false;
        p += r;
        n -= (size_t)r;
        offset += r;
    }
    return true;
}

bool valid_extents(const int64_t shape[4]) {
    __int128 total = 1;
    for (int k = 0; k < 4; ++k) {
        if (shape[k] <= 0) return false;
        total *= shape[k];
        if (total > ((__int128)1 << 60)) return false;
    }
    return true;
}

}  // namespace

int rows_data(const cubo_hip_cube *c, int64_t row0, int64_t nrows, const void **out) {
    const size_t row_bytes = (size_t)c->shape[1] * c->ndet * element_size(c->dtype);
    if (!c->streamed) {
        *out = (const char *)c->data + (size_t)row0 * row_bytes;
        return CUBO_HIP_OK;
    }
    const Feed *f = c->feed;
    for (int s = 0; f && s < 2; ++s)
        if (f->ring_row0[s] >= 0 && row0 >= f->ring_row0[s] && row0 + nrows <= f->ring_row0[s] + f->ring_rows[s]) {
            *out = (const char *)f->ring[s] + (size_t)(row0 - f->ring_row0[s]) * row_bytes;
            return CUBO_HIP_OK;
        }
    return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "scan rows [%lld, %lld) have not been fed to this streamed cube",
                (long long)row0, (long long)(row0 + nrows));
}

void free_feed(cubo_hip_cube *c) {
    Feed *f = c->feed;
    if (!f) return;
    DeviceGuard g(c->device);
    if (f->copy) (void)hipStreamSynchronize(f->copy);
    for (int s = 0; s < 2; ++s) {
        if (f->pinned[s]) (void)hipHostFree(f->pinned[s]);
        if (f->ring[s]) (void)hipFree(f->ring[s]);
        if (f->copied[s]) (void)hipEventDestroy(f->copied[s]);
        if (f->consumed[s]) (void)hipEventDestroy(f->consumed[s]);
    }
    if (f->copy) (void)hipStreamDestroy(f->copy);
    delete f;
    c->feed = nullptr;
}

}  // namespace cubo

using cubo::Fd;

extern "C" {

int cubo_hip_npy_read_header(const char *path, cubo_hip_npy_header *h) {
    if (!path || !h) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    cubo::Fd fd;
    int64_t size = 0;
    if (int st = cubo::open_file(path, &fd, &size)) return st;
    unsigned char pre[12] = {0};
    const ssize_t got = pread(fd.fd, pre, 12, 0);
    static const unsigned char magic[6] = {0x93, 'N', 'U', 'M', 'P', 'Y'};
    if (got < 10 || memcmp(pre, magic, 6) != 0) return cubo::not_npy("no NUMPY magic");
    const int major = pre[6], minor = pre[7];
    int64_t hlen = 0, hstart = 0;
    if (major == 1) {
        hlen = (int64_t)pre[8] | (int64_t)pre[9] << 8;
        hstart = 10;
    } else if (major == 2 || major == 3) {
        if (got < 12) return fail(CUBO_HIP_ERR_FILE, "notNPY: truncated version %d preamble", major);
        hlen = (int64_t)pre[8] | (int64_t)pre[9] << 8 | (int64_t)pre[10] << 16 | (int64_t)pre[11] << 24;
        hstart = 12;
    } else {
        return fail(CUBO_HIP_ERR_FILE, "notNPY: version %d.%d", major, minor);
    }
    if (hlen <= 0 || hlen >= ((int64_t)1 << 26)) return fail(CUBO_HIP_ERR_FILE, "notNPY: header length %lld", (long long)hlen);
    std::string text((size_t)hlen, '\0');
    if (!cubo::read_at(fd.fd, &text[0], (size_t)hlen, hstart)) return cubo::not_npy("truncated header");
    cubo::Fields fields;
    if (int st = cubo::parse_dict(text, &fields)) return st;
    if (!fields.has_descr) return cubo::not_npy("no 'descr'");
    if (!fields.has_fortran) return cubo::not_npy("no 'fortran_order'");
    if (!fields.has_shape) return cubo::not_npy("no 'shape'");
    if (fields.fortran) return fail(CUBO_HIP_ERR_FILE, "fortranOrderNotSupported");
    int t = 0;
    if (int st = cubo::element_type(fields.descr, &t)) return st;
    for (int64_t e : fields.shape)
        if (e < 0) return cubo::not_npy("negative extent in 'shape'");
    // the element and byte counts must fit a signed 64-bit integer
    __int128 bytes = cubo::element_bytes(t);
    for (int64_t e : fields.shape) {
        bytes *= e;
        if (bytes > (__int128)INT64_MAX) return cubo::not_npy("'shape' overflows the byte count");
    }
    if (fields.shape.size() > 32) return cubo::not_npy("more than 32 dimensions");
    memset(h, 0, sizeof *h);
    h->major_version = major;
    h->minor_version = minor;
    h->element_type = t;
    h->rank = (int32_t)fields.shape.size();
    for (size_t k = 0; k < fields.shape.size(); ++k) h->shape[k] = fields.shape[k];
    h->data_offset = hstart + hlen;
    return CUBO_HIP_OK;
}

int cubo_hip_cube_load_file(int device, const char *path, int64_t data_offset, int32_t element_type,
                            const int64_t shape[4], cubo_hip_cube **out) {
    if (!path || !shape || !out) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    const int eb = cubo::element_bytes(element_type);
    if (!eb) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "unknown element type %d", element_type);
    // MappedCube.init (MappedCube.swift:98-127): offset, open, byte count, size
    if (data_offset < 0 || data_offset % eb != 0)
        return fail(CUBO_HIP_ERR_FILE, "misalignedOffset: %lld for %d-byte elements", (long long)data_offset, eb);
    cubo::Fd fd;
    int64_t size = 0;
    if (int st = cubo::open_file(path, &fd, &size)) return st;
    if (!cubo::valid_extents(shape)) return fail(CUBO_HIP_ERR_FILE, "invalidShape");
    const __int128 cube_bytes = (__int128)shape[0] * shape[1] * shape[2] * shape[3] * eb;
    const __int128 needed = cube_bytes + data_offset;
    if (needed > (__int128)size)
        return fail(CUBO_HIP_ERR_FILE, "fileTooShort: %s needs %lld bytes, holds %lld", path, (long long)needed,
                    (long long)size);
    cubo_hip_cube *c = nullptr;
    if (int st = cubo_hip_cube_create(device, shape, cubo::held_as(element_type), &c)) return st;
    if (int st = cubo::make_feed(c)) { cubo_hip_cube_destroy(c); return st; }
    DeviceGuard g(device);
    const int threads = cubo::reader_threads();
    const int fdn = fd.fd;
    int st = cubo::feed_bytes(c, c->data, (size_t)cube_bytes, element_type, nullptr,
                              [&](void *pinned, size_t off, size_t len) {
        const bool ok = cubo::parallel_pieces(len, threads, [&](size_t o, size_t n) {
            return cubo::read_at(fdn, (char *)pinned + o, n, data_offset + (int64_t)(off + o));
        });
        if (!ok) fail(CUBO_HIP_ERR_FILE, "cannotOpen: %s: read failed at byte %lld", path, (long long)(data_offset + off));
        return ok;
    });
    if (st == CUBO_HIP_OK) {
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) st = hip_fail(e, "loading the cube");
    }
    cubo::free_feed(c);     // a loaded cube is resident: its pinned buffers go back
    if (st != CUBO_HIP_OK) { cubo_hip_cube_destroy(c); return st; }
    *out = c;
    return CUBO_HIP_OK;
}

int cubo_hip_cube_load_npy(int device, const char *path, cubo_hip_cube **out) {
    if (!path || !out) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    cubo_hip_npy_header h;
    if (int st = cubo_hip_npy_read_header(path, &h)) return st;
    if (h.rank != 4) return fail(CUBO_HIP_ERR_FILE, "wrongRank: expected 4, received %d", h.rank);
    return cubo_hip_cube_load_file(device, path, h.data_offset, h.element_type, h.shape, out);
}

int cubo_hip_cube_create_streamed(int device, const int64_t shape[4], int dtype, int64_t tile_rows,
                                  cubo_hip_cube **out) {
    if (!shape || !out) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return fail(CUBO_HIP_ERR_NO_DEVICE, "no HIP device is available; run the CPU path");
    if (device < 0 || device >= n) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "device %d of %d", device, n);
    if (!cubo::valid_extents(shape)) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "invalid cube shape");
    const int es = element_size(dtype);
    if (!es) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "unknown dtype %d", dtype);
    if (tile_rows < 1 || tile_rows > shape[0])
        return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "tile_rows %lld outside [1, %lld]", (long long)tile_rows,
                    (long long)shape[0]);
    DeviceGuard g(device);
    auto *c = new (std::nothrow) cubo_hip_cube();
    if (!c) return fail(CUBO_HIP_ERR_OUT_OF_MEMORY, "host allocation");
    c->device = device;
    memcpy(c->shape, shape, sizeof c->shape);
    c->dtype = dtype;
    c->npos = shape[0] * shape[1];
    c->ndet = shape[2] * shape[3];
    c->bytes = (uint64_t)c->npos * c->ndet * es;
    c->data = nullptr;
    c->streamed = true;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete c; return hip_fail(e, "hipStreamCreate"); }
    if (int st = cubo::make_feed(c)) { cubo_hip_cube_destroy(c); return st; }
    const size_t slot = (size_t)tile_rows * shape[1] * c->ndet * es;
    for (int s = 0; s < 2 && e == hipSuccess; ++s) e = hipMalloc(&c->feed->ring[s], slot);
    if (e != hipSuccess) { cubo_hip_cube_destroy(c); return hip_fail(e, "allocating the tile ring"); }
    c->feed->tile_rows = tile_rows;
    *out = c;
    return CUBO_HIP_OK;
}

int cubo_hip_cube_feed_rows(cubo_hip_cube *c, int64_t row0, int64_t nrows, const void *host, int32_t element_type) {
    if (!c || (!host && nrows > 0)) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (int st = cubo::rows_range(c, row0, nrows)) return st;
    const int eb = cubo::element_bytes(element_type);
    if (!eb) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "unknown element type %d", element_type);
    if (cubo::held_as(element_type) != c->dtype)
        return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "element type %d is not held as this cube's type", element_type);
    if (nrows == 0) return CUBO_HIP_OK;
    if (int st = cubo::make_feed(c)) return st;
    cubo::Feed *f = c->feed;
    DeviceGuard g(c->device);
    const size_t raw_row = (size_t)c->shape[1] * c->ndet * eb;
    void *dst = nullptr;
    hipEvent_t after = nullptr;
    int slot = -1;
    if (c->streamed) {
        if (nrows > f->tile_rows)
            return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "%lld rows exceed the ring's %lld-row slots", (long long)nrows,
                        (long long)f->tile_rows);
        slot = f->ring_next;
        f->ring_next ^= 1;
        // every launch that reads the slot's previous tile is already on the cube's stream
        CUBO_HIP_TRY(hipEventRecord(f->consumed[slot], c->stream), "hipEventRecord");
        after = f->consumed[slot];
        dst = f->ring[slot];
        f->ring_row0[slot] = -1;
    } else {
        ++c->generation;
        dst = (char *)c->data + (size_t)row0 * c->shape[1] * c->ndet * element_size(c->dtype);
        // earlier launches on the cube may still read these rows
        CUBO_HIP_TRY(hipEventRecord(f->consumed[0], c->stream), "hipEventRecord");
        after = f->consumed[0];
    }
    const int threads = cubo::reader_threads();
    const char *src = (const char *)host;
    int st = cubo::feed_bytes(c, dst, (size_t)nrows * raw_row, element_type, after,
                              [&](void *pinned, size_t off, size_t len) {
        cubo::parallel_pieces(len, threads, [&](size_t o, size_t n) {
            memcpy((char *)pinned + o, src + off + o, n);
            return true;
        });
        return true;
    });
    if (st) return st;
    if (c->streamed) {
        f->ring_row0[slot] = row0;
        f->ring_rows[slot] = nrows;
    }
    return CUBO_HIP_OK;
}

}  // extern "C"

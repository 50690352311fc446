// TEXMEX vector files (.fvecs / .ivecs / .bvecs) and recall: the data side of
// the reference's Dataset (include/dataset.h:14-204) behind the C ABI.
//
// A file is a sequence of records [int32 dim][dim elements of elem_bytes]. A
// read strips the headers (checking every one against the first) and lands
// the rows packed [count][dim], in host memory or streamed to device memory
// through two pinned staging buffers (disk read of chunk i+1 overlaps the
// host-to-device copy of chunk i). Writing (the reference's TODO
// saveXvecsData) and a recall that leaves its inputs untouched (the reference
// sorts the caller's arrays in place, dataset.h:137) complete it.
#include <hip/hip_runtime.h>

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "rplsh.h"

namespace rplsh {
namespace {

struct Fd {
  int fd = -1;
  ~Fd() {
    if (fd >= 0) ::close(fd);
  }
};

void read_exact(int fd, void* buf, size_t bytes, int64_t off, const char* path) {
  char* p = static_cast<char*>(buf);
  while (bytes > 0) {
    const ssize_t r = ::pread(fd, p, bytes, (off_t)off);
    if (r < 0 && errno == EINTR) continue;
    if (r <= 0) fail(RPLSH_ERR_INVALID_ARG, "%s: read failed at offset %lld", path, (long long)off);
    p += r;
    off += r;
    bytes -= (size_t)r;
  }
}

struct XvecsInfo {
  int64_t num = 0;
  int dim = 0;
  int64_t record = 0;  // bytes per record
};

XvecsInfo open_info(const char* path, int elem_bytes, Fd& f) {
  RPLSH_REQUIRE(path, "NULL path");
  RPLSH_REQUIRE(elem_bytes == 1 || elem_bytes == 4, "elem_bytes must be 4 (fvecs/ivecs) or 1 (bvecs)");
  f.fd = ::open(path, O_RDONLY);
  if (f.fd < 0) fail(RPLSH_ERR_INVALID_ARG, "%s: cannot open (%s)", path, std::strerror(errno));
  struct stat st {};
  if (::fstat(f.fd, &st) != 0) fail(RPLSH_ERR_INVALID_ARG, "%s: cannot stat", path);
  XvecsInfo info;
  const int64_t size = (int64_t)st.st_size;
  if (size == 0) return info;
  RPLSH_REQUIRE(size >= 4, "%s: truncated header", path);
  int32_t dim = 0;
  read_exact(f.fd, &dim, 4, 0, path);
  RPLSH_REQUIRE(dim > 0 && dim <= (1 << 20), "%s: bad dimension %d", path, dim);
  info.dim = dim;
  info.record = 4 + (int64_t)dim * elem_bytes;
  RPLSH_REQUIRE(size % info.record == 0, "%s: size %lld is not a whole number of %lld-byte records", path,
                (long long)size, (long long)info.record);
  info.num = size / info.record;
  return info;
}

// Strips the headers of `count` records in raw (in place: data moves left)
// and checks each against dim. Returns the packed byte count.
size_t compact(char* raw, int64_t count, int dim, int elem_bytes, int64_t first, const char* path) {
  const size_t row = (size_t)dim * elem_bytes;
  for (int64_t i = 0; i < count; ++i) {
    int32_t h;
    std::memcpy(&h, raw + i * (row + 4), 4);
    if (h != dim)
      fail(RPLSH_ERR_INVALID_ARG, "%s: record %lld has dimension %d, expected %d", path,
           (long long)(first + i), h, dim);
    std::memmove(raw + i * row, raw + i * (row + 4) + 4, row);
  }
  return (size_t)count * row;
}

}  // namespace
}  // namespace rplsh

using namespace rplsh;

extern "C" {

rplsh_status rplsh_xvecs_info(const char* path, int elem_bytes, int64_t* num, int* dim) {
  return guard([&] {
    RPLSH_REQUIRE(num && dim, "NULL output");
    Fd f;
    const XvecsInfo info = open_info(path, elem_bytes, f);
    *num = info.num;
    *dim = info.dim;
  });
}

rplsh_status rplsh_xvecs_read(const char* path, int elem_bytes, int64_t first, int64_t count, void* out,
                              int on_device, int device) {
  return guard([&] {
    Fd f;
    const XvecsInfo info = open_info(path, elem_bytes, f);
    RPLSH_REQUIRE(first >= 0 && count >= 0 && first + count <= info.num,
                  "%s: records [%lld, %lld) outside [0, %lld)", path, (long long)first,
                  (long long)(first + count), (long long)info.num);
    if (count == 0) return;
    RPLSH_REQUIRE(out, "NULL output");
    const size_t row = (size_t)info.dim * elem_bytes;
    if (!on_device) {
      // Host: read raw records into the tail of a scratch buffer chunk by chunk.
      const int64_t per = std::max<int64_t>(1, (64ll << 20) / info.record);
      std::vector<char> raw((size_t)std::min(per, count) * info.record);
      char* dst = static_cast<char*>(out);
      for (int64_t r = 0; r < count; r += per) {
        const int64_t c = std::min(per, count - r);
        read_exact(f.fd, raw.data(), (size_t)(c * info.record), (first + r) * info.record, path);
        const size_t b = compact(raw.data(), c, info.dim, elem_bytes, first + r, path);
        std::memcpy(dst, raw.data(), b);
        dst += b;
      }
      return;
    }
    DeviceScope scope(device);
    const int64_t per = std::max<int64_t>(1, (64ll << 20) / info.record);
    const size_t stage_bytes = (size_t)std::min(per, count) * info.record;
    void* stage[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    hipStream_t st = nullptr;
    auto cleanup = [&] {
      if (st) (void)hipStreamSynchronize(st);
      for (int i = 0; i < 2; ++i) {
        if (done[i]) (void)hipEventDestroy(done[i]);
        if (stage[i]) (void)hipHostFree(stage[i]);
      }
      if (st) (void)hipStreamDestroy(st);
    };
    try {
      RPLSH_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
      for (int i = 0; i < 2; ++i) {
        RPLSH_HIP(hipHostMalloc(&stage[i], stage_bytes, hipHostMallocDefault));
        RPLSH_HIP(hipEventCreateWithFlags(&done[i], hipEventDisableTiming));
      }
      char* dst = static_cast<char*>(out);
      int buf = 0;
      bool used[2] = {false, false};
      for (int64_t r = 0; r < count; r += per, buf ^= 1) {
        const int64_t c = std::min(per, count - r);
        if (used[buf]) RPLSH_HIP(hipEventSynchronize(done[buf]));  // its previous copy is done
        char* s = static_cast<char*>(stage[buf]);
        read_exact(f.fd, s, (size_t)(c * info.record), (first + r) * info.record, path);
        const size_t b = compact(s, c, info.dim, elem_bytes, first + r, path);
        RPLSH_HIP(hipMemcpyAsync(dst, s, b, hipMemcpyHostToDevice, st));
        RPLSH_HIP(hipEventRecord(done[buf], st));
        used[buf] = true;
        dst += b;
      }
      RPLSH_HIP(hipStreamSynchronize(st));
    } catch (...) {
      cleanup();
      throw;
    }
    cleanup();
    (void)row;
  });
}

rplsh_status rplsh_xvecs_write(const char* path, int elem_bytes, const void* data, int64_t num, int dim,
                               int append) {
  return guard([&] {
    RPLSH_REQUIRE(path, "NULL path");
    RPLSH_REQUIRE(elem_bytes == 1 || elem_bytes == 4, "elem_bytes must be 4 or 1");
    RPLSH_REQUIRE(dim > 0 && num >= 0, "bad shape");
    RPLSH_REQUIRE(num == 0 || data, "NULL data");
    FILE* fp = std::fopen(path, append ? "ab" : "wb");
    if (!fp) fail(RPLSH_ERR_INVALID_ARG, "%s: cannot open for writing (%s)", path, std::strerror(errno));
    const size_t row = (size_t)dim * elem_bytes;
    const char* p = static_cast<const char*>(data);
    bool ok = true;
    const int32_t h = dim;
    for (int64_t i = 0; i < num && ok; ++i) {
      ok = std::fwrite(&h, 4, 1, fp) == 1 && std::fwrite(p + i * row, 1, row, fp) == row;
    }
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok) fail(RPLSH_ERR_INVALID_ARG, "%s: write failed", path);
  });
}

rplsh_status rplsh_recall_at(const int32_t* result, int64_t nq, int ld_result, const int32_t* truth,
                             int ld_truth, int k, int32_t* missed, double* recall) {
  return guard([&] {
    RPLSH_REQUIRE(recall, "NULL output");
    RPLSH_REQUIRE(k > 0 && k <= ld_result && k <= ld_truth, "k must be in [1, min(ld_result, ld_truth)]");
    RPLSH_REQUIRE(nq >= 0 && (nq == 0 || (result && truth)), "NULL input");
    std::vector<int32_t> a((size_t)k), b((size_t)k);
    double sum = 0.0;
    for (int64_t q = 0; q < nq; ++q) {
      std::copy(result + q * ld_result, result + q * ld_result + k, a.begin());
      std::copy(truth + q * ld_truth, truth + q * ld_truth + k, b.begin());
      std::sort(a.begin(), a.end());
      std::sort(b.begin(), b.end());
      int hit = 0;
      size_t i = 0, j = 0;
      while (i < a.size() && j < b.size()) {
        if (a[i] < b[j]) {
          ++i;
        } else if (b[j] < a[i]) {
          ++j;
        } else {
          if (a[i] >= 0) ++hit;  // -1 pads never count
          ++i;
          ++j;
        }
      }
      sum += (double)hit / k;
      if (missed) missed[q] = k - hit;
    }
    *recall = nq > 0 ? sum / (double)nq : 0.0;
  });
}

}  // extern "C"

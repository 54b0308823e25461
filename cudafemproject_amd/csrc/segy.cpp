// segy.cpp -- SEG-Y output of receiver traces (femgpu_write_segy in femgpu.h).
//
// The layout the reference writes through its vendored SEG-Y writer
// (convertToSEGY, VTKfile/src/VTKfile.cpp:89-138, with CSegyOut of
// thirdparty/segyio: setSegyFormat(5), setSegyID(0)), as in the reference's
// configs/test.sgy: a zero 3200-byte textual header; a 400-byte binary
// header holding only the trace count (bytes 3213-3214), the sample interval
// in microseconds (3217-3218), the samples per trace (3221-3222) and the
// format code 5 (3225-3226); per trace a 240-byte header holding the id
// (bytes 9-12), the sample count (115-116) and the interval (117-118),
// followed by the samples as 4-byte IEEE floats.  Everything big-endian;
// 16-bit fields keep the low 16 bits of their value.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "femgpu.h"

namespace {

void put16(std::vector<unsigned char> &b, size_t off, int v) {
  const uint16_t u = uint16_t(uint32_t(v) & 0xffffu);
  b[off] = uint8_t(u >> 8);
  b[off + 1] = uint8_t(u & 0xff);
}

void put32(std::vector<unsigned char> &b, size_t off, uint32_t u) {
  b[off] = uint8_t(u >> 24);
  b[off + 1] = uint8_t((u >> 16) & 0xff);
  b[off + 2] = uint8_t((u >> 8) & 0xff);
  b[off + 3] = uint8_t(u & 0xff);
}

}  // namespace

extern "C" int femgpu_write_segy(const char *path, const float *traces, int n_traces,
                                 int n_samples, float sample_interval, int trace_id) {
  if (!path || (!traces && n_traces * n_samples > 0) || n_traces < 0 || n_samples < 0)
    return FEMGPU_E_ARG;
  // interval in microseconds, formed in float as the writer does
  const int interval_us = int(sample_interval * float(1000000L));
  std::vector<unsigned char> head(3600, 0);
  put16(head, 3212, n_traces);
  put16(head, 3216, interval_us);
  put16(head, 3220, n_samples);
  put16(head, 3224, 5);
  FILE *f = std::fopen(path, "wb");
  if (!f) return FEMGPU_E_ARG;
  bool ok = std::fwrite(head.data(), 1, head.size(), f) == head.size();
  std::vector<unsigned char> th(240), data(size_t(n_samples) * 4);
  for (int t = 0; t < n_traces && ok; ++t) {
    std::fill(th.begin(), th.end(), 0);
    put32(th, 8, uint32_t(trace_id));
    put16(th, 114, n_samples);
    put16(th, 116, interval_us);
    for (int s = 0; s < n_samples; ++s) {
      uint32_t u;
      const float v = traces[size_t(t) * n_samples + s];
      std::memcpy(&u, &v, 4);
      put32(data, size_t(s) * 4, u);
    }
    ok = std::fwrite(th.data(), 1, th.size(), f) == th.size() &&
         std::fwrite(data.data(), 1, data.size(), f) == data.size();
  }
  ok = (std::fclose(f) == 0) && ok;
  return ok ? FEMGPU_OK : FEMGPU_E_ARG;
}

// jms_io.cpp -- EMS/SLI slice files and .bin STEM result channels
// (include/jms_io.h).
#include "jms_io.h"

#include <cstdio>
#include <cstring>
#include <vector>

namespace {

// header field offsets (prm_slice_file_info.h:37-55)
constexpr long kPosDataSizeX = 0x0000;
constexpr long kPosDataSizeY = 0x0004;
constexpr long kPosTitle = 0x004C;
constexpr long kPosExtHeaderVer = 0x0074;
constexpr long kPosVariantNum = 0x0078;
constexpr long kPosContentType = 0x0080;
constexpr long kPosAltOffset = 0x0088;
constexpr long kPosItabOffset = 0x008C;
constexpr long kPosThickness = 0x1034;
constexpr long kPosHighTension = 0x1040;
constexpr long kPosPhysSizeX = 0x1044;
constexpr long kPosPhysSizeY = 0x1048;
constexpr long kPosItab = 0x1100;
constexpr long kPosData = 0x2000;
constexpr uint32_t kHeaderVersion = 2020052801u;  // SLICEPARAMS_EXTHDRVER (.h:32)

template <class T>
T swap_bytes(T v) {
  unsigned char b[sizeof(T)];
  std::memcpy(b, &v, sizeof(T));
  for (size_t i = 0; i < sizeof(T) / 2; ++i) std::swap(b[i], b[sizeof(T) - 1 - i]);
  std::memcpy(&v, b, sizeof(T));
  return v;
}

template <class T>
bool read_at(FILE* f, long pos, T* v, bool swap) {
  if (std::fseek(f, pos, SEEK_SET) != 0) return false;
  if (std::fread(v, sizeof(T), 1, f) != 1) return false;
  if (swap) *v = swap_bytes(*v);
  return true;
}

}  // namespace

extern "C" int jms_ems_read_header(const char* path, jms_ems_header* h) {
  if (!path || !h) return 1;
  FILE* f = std::fopen(path, "rb");
  if (!f) return 2;
  std::memset(h, 0, sizeof(*h));
  int32_t ichk = 0;
  bool ok = read_at(f, 0, &ichk, false);
  const bool swap = ok && (ichk < 0 || ichk > 16384);  // check_ems_endian
  h->byte_swapped = swap ? 1 : 0;
  h->data_offset = (uint64_t)kPosData;
  h->structure_offset = (uint64_t)kPosItab;
  ok = ok && read_at(f, kPosDataSizeX, &h->grid_x, swap) && read_at(f, kPosDataSizeY, &h->grid_y, swap);
  if (ok) {
    ok = std::fseek(f, kPosTitle, SEEK_SET) == 0 && std::fread(h->title, 1, sizeof(h->title), f) == sizeof(h->title);
  }
  ok = ok && read_at(f, kPosExtHeaderVer, &h->version, swap) && read_at(f, kPosContentType, &h->data_type, swap);
  if (ok) {
    uint32_t u = 0;
    switch (h->version) {
      case 2010112301u:
        ok = read_at(f, kPosVariantNum, &h->num_var, swap);
        break;
      case 2012071101u:
      case 2020052801u:
        ok = read_at(f, kPosVariantNum, &h->num_var, swap);
        if (ok && (ok = read_at(f, kPosAltOffset, &u, swap))) h->data_offset = u;
        if (ok && (ok = read_at(f, kPosItabOffset, &u, swap))) h->structure_offset = u;
        break;
      default:  // no extended header: one variant
        h->num_var = 1;
    }
  }
  ok = ok && read_at(f, kPosThickness, &h->sz, swap) && read_at(f, kPosHighTension, &h->ekv, swap) &&
       read_at(f, kPosPhysSizeX, &h->sx, swap) && read_at(f, kPosPhysSizeY, &h->sy, swap);
  std::fclose(f);
  if (!ok) return 3;
  // grid edges lie in [1, 16384] (the byte-order test relies on it); at
  // least one variant
  if (h->grid_x < 1 || h->grid_x > 16384 || h->grid_y < 1 || h->grid_y > 16384 || h->num_var < 1) return 4;
  return 0;
}

extern "C" int jms_ems_read_data(const char* path, const jms_ems_header* h, float* data) {
  if (!path || !h || !data) return 1;
  const size_t nf = (size_t)2 * h->grid_x * h->grid_y * h->num_var;
  if (nf == 0 || h->data_offset < (uint64_t)kPosData) return 2;  // invalid header presets
  FILE* f = std::fopen(path, "rb");
  if (!f) return 3;
  bool ok = std::fseek(f, (long)h->data_offset, SEEK_SET) == 0 && std::fread(data, sizeof(float), nf, f) == nf;
  std::fclose(f);
  return ok ? 0 : 4;
}

extern "C" int jms_ems_write(const char* path, const jms_ems_header* h, const float* data) {
  if (!path || !h || !data) return 1;
  const uint64_t doff = h->data_offset ? h->data_offset : (uint64_t)kPosData;
  const uint64_t soff = h->structure_offset ? h->structure_offset : (uint64_t)kPosItab;
  if (doff < (uint64_t)kPosData || soff < (uint64_t)kPosItab || soff + 12 > doff) return 2;
  std::vector<unsigned char> buf((size_t)doff, 0);
  auto put = [&](long pos, const void* v, size_t n) { std::memcpy(&buf[(size_t)pos], v, n); };
  const uint32_t ver = kHeaderVersion, d32 = (uint32_t)doff, s32 = (uint32_t)soff;
  put(kPosDataSizeX, &h->grid_x, 4);
  put(kPosDataSizeY, &h->grid_y, 4);
  put(kPosTitle, h->title, sizeof(h->title));
  put(kPosExtHeaderVer, &ver, 4);
  put(kPosVariantNum, &h->num_var, 4);
  put(kPosContentType, &h->data_type, 4);
  put(kPosAltOffset, &d32, 4);
  put(kPosItabOffset, &s32, 4);
  put(kPosThickness, &h->sz, 4);
  put(kPosHighTension, &h->ekv, 4);
  put(kPosPhysSizeX, &h->sx, 4);
  put(kPosPhysSizeY, &h->sy, 4);
  // the atom table at soff stays empty: 0 types, 0 positions, 0 transitions
  FILE* f = std::fopen(path, "wb");
  if (!f) return 3;
  const size_t nf = (size_t)2 * h->grid_x * h->grid_y * h->num_var;
  bool ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size() &&
            std::fwrite(data, sizeof(float), nf, f) == nf;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? 0 : 4;
}

extern "C" int jms_stem_images(int npos, int ndetslc, int ndet, const float* res, float weight, float* out) {
  if (npos < 1 || ndetslc < 1 || ndet < 1 || !res || !out || weight == 0.f) return 1;
  for (int j = 0; j < ndet; ++j)
    for (int i = 0; i < ndetslc; ++i)
      for (int p = 0; p < npos; ++p)
        out[(size_t)p + (size_t)npos * i + (size_t)ndetslc * npos * j] =
            res[((size_t)p * ndetslc + i) * ndet + j] / weight;
  return 0;
}

extern "C" int jms_bin_write_channel(const char* path, const float* images, int npos, int ndetslc, int ndet,
                                     int idet) {
  if (!path || !images || npos < 1 || ndetslc < 1 || idet < 0 || idet >= ndet) return 1;
  FILE* f = std::fopen(path, "wb");
  if (!f) return 2;
  const size_t n = (size_t)npos * ndetslc;
  bool ok = std::fwrite(images + n * idet, sizeof(float), n, f) == n;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? 0 : 3;
}

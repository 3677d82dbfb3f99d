// Minimal 8-bit RGB PNG writer (zlib "stored" deflate blocks, no
// compression library needed). The reference saves its film through CImg.
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../../include/muonx.h"

namespace {

uint32_t Crc32(const uint8_t *p, size_t n, uint32_t crc = 0xffffffffu) {
  static uint32_t table[256];
  static bool init = false;
  if (!init) {
    for (uint32_t i = 0; i < 256; ++i) {
      uint32_t c = i;
      for (int k = 0; k < 8; ++k) c = (c & 1) ? 0xedb88320u ^ (c >> 1) : c >> 1;
      table[i] = c;
    }
    init = true;
  }
  for (size_t i = 0; i < n; ++i) crc = table[(crc ^ p[i]) & 0xff] ^ (crc >> 8);
  return crc;
}

void Put32(std::vector<uint8_t> &v, uint32_t x) {
  v.push_back(uint8_t(x >> 24));
  v.push_back(uint8_t(x >> 16));
  v.push_back(uint8_t(x >> 8));
  v.push_back(uint8_t(x));
}

void Chunk(std::vector<uint8_t> &out, const char *type, const std::vector<uint8_t> &data) {
  Put32(out, uint32_t(data.size()));
  std::vector<uint8_t> td(type, type + 4);
  td.insert(td.end(), data.begin(), data.end());
  out.insert(out.end(), td.begin(), td.end());
  Put32(out, Crc32(td.data(), td.size()) ^ 0xffffffffu);
}

}  // namespace

extern "C" int muonx_write_png(const char *path, const uint8_t *rgb8, int32_t w, int32_t h) {
  if (!path || !rgb8 || w <= 0 || h <= 0) return MUONX_ERR_INVALID;
  std::vector<uint8_t> raw;
  raw.reserve(size_t(h) * (size_t(w) * 3 + 1));
  for (int32_t y = 0; y < h; ++y) {
    raw.push_back(0);  // filter: none
    raw.insert(raw.end(), rgb8 + size_t(y) * w * 3, rgb8 + size_t(y + 1) * w * 3);
  }
  std::vector<uint8_t> z{0x78, 0x01};
  uint32_t a = 1, b = 0;
  for (uint8_t c : raw) {
    a = (a + c) % 65521u;
    b = (b + a) % 65521u;
  }
  size_t pos = 0;
  do {
    const size_t len = std::min<size_t>(65535, raw.size() - pos);
    const bool last = pos + len == raw.size();
    z.push_back(last ? 1 : 0);
    z.push_back(uint8_t(len));
    z.push_back(uint8_t(len >> 8));
    z.push_back(uint8_t(~len));
    z.push_back(uint8_t((~len) >> 8));
    z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + len);
    pos += len;
  } while (pos < raw.size());
  Put32(z, (b << 16) | a);

  std::vector<uint8_t> png{0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
  std::vector<uint8_t> ihdr;
  Put32(ihdr, uint32_t(w));
  Put32(ihdr, uint32_t(h));
  ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});
  Chunk(png, "IHDR", ihdr);
  Chunk(png, "IDAT", z);
  Chunk(png, "IEND", {});
  FILE *f = std::fopen(path, "wb");
  if (!f) return MUONX_ERR_IO;
  const bool ok = std::fwrite(png.data(), 1, png.size(), f) == png.size();
  std::fclose(f);
  return ok ? MUONX_OK : MUONX_ERR_IO;
}

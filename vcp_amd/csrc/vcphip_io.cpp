// Matrix files of VCP: the portable ".vcpmat" format (vcp/vcp_portable_fio.hpp,
// docs/file_io.md) and the older native ".matrix_*" format (vcp/vcp_fio.hpp),
// the formats in which VCP programs store the matrices that the verified
// products consume.  Host code only.
//
// Portable: magic "VCPMAT2\0"; little-endian u32 version (2), type id,
// layout (1 = column-major), reserved (0); u64 rows, columns, payload bytes;
// the payload in column-major order (int32, or binary64 words: interval
// {lower, upper}, dd {a1, a2}, interval<dd> {lower.a1, lower.a2, upper.a1,
// upper.a2}); u64 CRC-64 (polynomial 0x42F0E1EBA9EA3693, MSB first, initial
// value 0, no final xor) over header and payload; nothing after it.
//
// Native: a type tag ("integer", "double", "interval_double",
// "kv_doubledouble", "interval_kv_dd") without terminator, int32 rows and
// columns, then the elements in row-major order (the writing host's byte
// order; little-endian here).
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/vcphip.h"

namespace {

const unsigned char kMagic[8] = {'V', 'C', 'P', 'M', 'A', 'T', '2', '\0'};
const uint32_t kVersion = 2, kColumnMajor = 1;
const char* const kTags[6] = {nullptr, "integer", "double", "interval_double", "kv_doubledouble", "interval_kv_dd"};

int elem_bytes(int type) {
  switch (type) {
    case VCPHIP_MAT_INT32: return 4;
    case VCPHIP_MAT_DOUBLE: return 8;
    case VCPHIP_MAT_INTERVAL:
    case VCPHIP_MAT_DD: return 16;
    case VCPHIP_MAT_INTERVAL_DD: return 32;
    default: return 0;
  }
}

void crc64_bytes(uint64_t& crc, const unsigned char* p, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    crc ^= (uint64_t)p[i] << 56;
    for (int b = 0; b < 8; ++b) crc = (crc & 0x8000000000000000ull) ? (crc << 1) ^ 0x42F0E1EBA9EA3693ull : crc << 1;
  }
}

void put_le(std::vector<unsigned char>& out, uint64_t v, int nbytes) {
  for (int i = 0; i < nbytes; ++i) out.push_back((unsigned char)(v >> (8 * i)));
}

uint64_t get_le(const unsigned char* p, int nbytes) {
  uint64_t v = 0;
  for (int i = 0; i < nbytes; ++i) v |= (uint64_t)p[i] << (8 * i);
  return v;
}

bool read_file(const char* path, std::vector<unsigned char>& buf) {
  FILE* f = fopen(path, "rb");
  if (!f) return false;
  buf.clear();
  unsigned char tmp[1 << 16];
  size_t got;
  while ((got = fread(tmp, 1, sizeof(tmp), f)) > 0) buf.insert(buf.end(), tmp, tmp + got);
  const bool ok = !ferror(f);
  fclose(f);
  return ok;
}

// write to path.tmp, then rename over path (an existing file survives a failed write)
bool write_file(const char* path, const std::vector<unsigned char>& buf) {
  const std::string tmp = std::string(path) + ".tmp";
  FILE* f = fopen(tmp.c_str(), "wb");
  if (!f) return false;
  bool ok = fwrite(buf.data(), 1, buf.size(), f) == buf.size();
  ok = (fclose(f) == 0) && ok;
  if (ok && rename(tmp.c_str(), path) != 0) ok = false;
  if (!ok) remove(tmp.c_str());
  return ok;
}

struct Parsed {
  int type = 0, rows = 0, cols = 0;
  bool portable = false;
  const unsigned char* payload = nullptr;   // portable: column-major; native: row-major
};

bool sizes_ok(uint64_t rows, uint64_t cols) {
  if ((rows == 0) != (cols == 0)) return false;
  if (rows > 0x7fffffffull || cols > 0x7fffffffull) return false;
  return rows == 0 || rows * cols <= 0x7fffffffull;
}

int parse(const std::vector<unsigned char>& b, Parsed& p) {
  if (b.size() >= 8 && memcmp(b.data(), kMagic, 8) == 0) {
    if (b.size() < 48 + 8) return VCPHIP_ERR_IO;
    if (get_le(&b[8], 4) != kVersion || get_le(&b[16], 4) != kColumnMajor || get_le(&b[20], 4) != 0)
      return VCPHIP_ERR_IO;
    const int type = (int)get_le(&b[12], 4);
    const uint64_t rows = get_le(&b[24], 8), cols = get_le(&b[32], 8), bytes = get_le(&b[40], 8);
    if (elem_bytes(type) == 0 || !sizes_ok(rows, cols)) return VCPHIP_ERR_IO;
    if (bytes != rows * cols * (uint64_t)elem_bytes(type) || b.size() != 48 + bytes + 8) return VCPHIP_ERR_IO;
    uint64_t crc = 0;
    crc64_bytes(crc, b.data(), 48 + bytes);
    if (crc != get_le(&b[48 + bytes], 8)) return VCPHIP_ERR_IO;
    p.type = type;
    p.rows = (int)rows;
    p.cols = (int)cols;
    p.portable = true;
    p.payload = b.data() + 48;
    return VCPHIP_OK;
  }
  // native format: the longest tag that the file starts with
  int type = 0;
  size_t tag_len = 0;
  for (int t = 1; t <= 5; ++t) {
    const size_t L = strlen(kTags[t]);
    if (b.size() >= L && memcmp(b.data(), kTags[t], L) == 0 && L > tag_len) {
      type = t;
      tag_len = L;
    }
  }
  if (!type || b.size() < tag_len + 8) return VCPHIP_ERR_IO;
  const int32_t rows = (int32_t)get_le(&b[tag_len], 4), cols = (int32_t)get_le(&b[tag_len + 4], 4);
  if (rows < 0 || cols < 0 || !sizes_ok((uint64_t)rows, (uint64_t)cols)) return VCPHIP_ERR_IO;
  if (b.size() != tag_len + 8 + (size_t)rows * cols * elem_bytes(type)) return VCPHIP_ERR_IO;
  p.type = type;
  p.rows = rows;
  p.cols = cols;
  p.portable = false;
  p.payload = b.data() + tag_len + 8;
  return VCPHIP_OK;
}

}  // namespace

extern "C" int vcphip_crc64(const void* data, size_t bytes, uint64_t* crc) {
  if ((!data && bytes) || !crc) return VCPHIP_ERR_ARGUMENT;
  uint64_t c = 0;
  crc64_bytes(c, (const unsigned char*)data, bytes);
  *crc = c;
  return VCPHIP_OK;
}

extern "C" int vcphip_mat_info(const char* path, int* type, int* rows, int* cols, int* portable) {
  if (!path || !type || !rows || !cols) return VCPHIP_ERR_ARGUMENT;
  std::vector<unsigned char> b;
  if (!read_file(path, b)) return VCPHIP_ERR_IO;
  Parsed p;
  const int rc = parse(b, p);
  if (rc) return rc;
  *type = p.type;
  *rows = p.rows;
  *cols = p.cols;
  if (portable) *portable = p.portable ? 1 : 0;
  return VCPHIP_OK;
}

extern "C" int vcphip_mat_read(const char* path, int type, int rows, int cols, void* data, int ld) {
  if (!path || (!data && rows > 0 && cols > 0)) return VCPHIP_ERR_ARGUMENT;
  if (elem_bytes(type) == 0) return VCPHIP_ERR_ARGUMENT;
  if (rows < 0 || cols < 0 || ld < (rows > 1 ? rows : 1)) return VCPHIP_ERR_DIMENSION;
  std::vector<unsigned char> b;
  if (!read_file(path, b)) return VCPHIP_ERR_IO;
  Parsed p;
  const int rc = parse(b, p);
  if (rc) return rc;
  if (p.type != type) return VCPHIP_ERR_IO;   // the reference's data type mismatch
  if (p.rows != rows || p.cols != cols) return VCPHIP_ERR_DIMENSION;
  const size_t eb = (size_t)elem_bytes(type);
  unsigned char* out = (unsigned char*)data;
  for (int j = 0; j < cols; ++j)
    for (int i = 0; i < rows; ++i) {
      const size_t src = p.portable ? ((size_t)i + (size_t)rows * j) : ((size_t)i * cols + j);
      const unsigned char* s = p.payload + src * eb;
      unsigned char* d = out + ((size_t)i + (size_t)ld * j) * eb;
      for (size_t w = 0; w < eb; w += 4) {   // little-endian words into host order
        const uint32_t v = (uint32_t)get_le(s + w, 4);
        memcpy(d + w, &v, 4);
      }
    }
  return VCPHIP_OK;
}

static int mat_write(const char* path, int type, int rows, int cols, const void* data, int ld, bool portable) {
  if (!path || (!data && rows > 0 && cols > 0)) return VCPHIP_ERR_ARGUMENT;
  if (elem_bytes(type) == 0) return VCPHIP_ERR_ARGUMENT;
  if (rows < 0 || cols < 0 || ld < (rows > 1 ? rows : 1)) return VCPHIP_ERR_DIMENSION;
  if (!sizes_ok((uint64_t)rows, (uint64_t)cols)) return VCPHIP_ERR_DIMENSION;
  const size_t eb = (size_t)elem_bytes(type), n = (size_t)rows * cols;
  std::vector<unsigned char> out;
  out.reserve(64 + n * eb);
  if (portable) {
    out.insert(out.end(), kMagic, kMagic + 8);
    put_le(out, kVersion, 4);
    put_le(out, (uint64_t)type, 4);
    put_le(out, kColumnMajor, 4);
    put_le(out, 0, 4);
    put_le(out, (uint64_t)rows, 8);
    put_le(out, (uint64_t)cols, 8);
    put_le(out, (uint64_t)(n * eb), 8);
  } else {
    out.insert(out.end(), kTags[type], kTags[type] + strlen(kTags[type]));
    put_le(out, (uint32_t)rows, 4);
    put_le(out, (uint32_t)cols, 4);
  }
  const unsigned char* in = (const unsigned char*)data;
  auto put_elem = [&](int i, int j) {
    const unsigned char* s = in + ((size_t)i + (size_t)ld * j) * eb;
    for (size_t w = 0; w < eb; w += 4) {
      uint32_t v;
      memcpy(&v, s + w, 4);
      put_le(out, v, 4);
    }
  };
  if (portable) {
    for (int j = 0; j < cols; ++j)
      for (int i = 0; i < rows; ++i) put_elem(i, j);
    uint64_t crc = 0;
    crc64_bytes(crc, out.data(), out.size());
    put_le(out, crc, 8);
  } else {
    for (int i = 0; i < rows; ++i)
      for (int j = 0; j < cols; ++j) put_elem(i, j);
  }
  return write_file(path, out) ? VCPHIP_OK : VCPHIP_ERR_IO;
}

extern "C" int vcphip_mat_write(const char* path, int type, int rows, int cols, const void* data, int ld,
                                int portable) {
  return mat_write(path, type, rows, cols, data, ld, portable != 0);
}

// G2S .bgrid files: the serialized DataImage, plain or gzip-compressed.
//
// Byte layout (include/DataImage.hpp:143-170 writes it, :172-224 checks it;
// every field is 4-byte aligned, host byte order):
//   size_t   fullSize          total bytes, this field included
//   unsigned nbDim             1..16
//   unsigned dims[nbDim]       each >= 1, dimension 0 first
//   unsigned nbVariable        1..4096
//   unsigned types[nbVariable] 0 continuous, 1 categorical
//   unsigned encoding          0 float, 1 integer, 2 unsigned integer
//   float    data[prod(dims) * nbVariable]   variable fastest, dim 0 next
// A buffer is valid when every field is in range and fullSize equals both
// the buffer size and the header + data size; at most 2^30 bytes. Files are
// read as gzip when they carry the gzip magic, else as plain bytes
// (src/DataImage.cpp:35-124, :126-163), with the same 2^30-byte cap.
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <zlib.h>

#include <algorithm>
#include <limits>
#include <string>
#include <vector>

#include "qsmi.h"

extern "C" int qsmi_internal_fail(int code, const char* msg);

namespace {

constexpr size_t kMaxBytes = size_t(1) << 30;
constexpr unsigned kMaxDims = 16;
constexpr unsigned kMaxVars = 4096;

int bfail(int code, const std::string& m) { return qsmi_internal_fail(code, ("bgrid: " + m).c_str()); }

// header bytes for ndim dimensions and nVar variables
size_t header_bytes(unsigned ndim, unsigned nVar) { return sizeof(size_t) + 4 * (1 + ndim + 1 + nVar + 1); }

uint32_t rd32(const unsigned char* p) {
  uint32_t v;
  memcpy(&v, p, 4);
  return v;
}

// Full check of a raw buffer; on success the header fields are returned.
bool parse(const unsigned char* raw, size_t size, unsigned* ndim, unsigned* dims, unsigned* nVar, size_t* cells,
           size_t* dataOff) {
  if (!raw || size < sizeof(size_t) + 12 || size > kMaxBytes) return false;
  size_t full;
  memcpy(&full, raw, sizeof(size_t));
  if (full != size) return false;
  size_t o = sizeof(size_t);
  if (o + 4 > size) return false;
  const unsigned nd = rd32(raw + o);
  o += 4;
  if (nd == 0 || nd > kMaxDims) return false;
  size_t c = 1;
  for (unsigned i = 0; i < nd; ++i) {
    if (o + 4 > size) return false;
    const unsigned d = rd32(raw + o);
    o += 4;
    if (d == 0 || c > std::numeric_limits<size_t>::max() / d) return false;
    c *= d;
    if (dims) dims[i] = d;
  }
  if (o + 4 > size) return false;
  const unsigned nv = rd32(raw + o);
  o += 4;
  if (nv == 0 || nv > kMaxVars || c > std::numeric_limits<size_t>::max() / nv) return false;
  for (unsigned v = 0; v < nv; ++v) {
    if (o + 4 > size) return false;
    const unsigned t = rd32(raw + o);
    o += 4;
    if (t > 1u) return false;
  }
  if (o + 4 > size) return false;
  const unsigned enc = rd32(raw + o);
  o += 4;
  if (enc > 2u) return false;
  const size_t n = c * nv;
  if (n > kMaxBytes / sizeof(float)) return false;
  if (!(o <= size && n * sizeof(float) == size - o)) return false;
  if (ndim) *ndim = nd;
  if (nVar) *nVar = nv;
  if (cells) *cells = c;
  if (dataOff) *dataOff = o;
  return true;
}

bool read_all_gz(const char* path, std::vector<unsigned char>& out) {
  gzFile f = gzopen(path, "rb");
  if (!f) return false;
  out.clear();
  unsigned char buf[1 << 16];
  bool ok = true;
  for (;;) {
    const int n = gzread(f, buf, sizeof(buf));
    if (n > 0) {
      if (out.size() + (size_t)n > kMaxBytes) {
        ok = false;
        break;
      }
      out.insert(out.end(), buf, buf + n);
      continue;
    }
    if (n < 0) ok = false;
    break;
  }
  gzclose(f);
  return ok;
}

}  // namespace

extern "C" {

int qsmi_bgrid_bytes(int ndim, const unsigned* dims, int nVar, size_t* bytes) {
  if (!dims || !bytes) return bfail(QSMI_ERR_ARG, "null argument");
  if (ndim < 1 || ndim > (int)kMaxDims || nVar < 1 || nVar > (int)kMaxVars)
    return bfail(QSMI_ERR_ARG, "dimension or variable count out of range");
  size_t c = (size_t)nVar;
  for (int i = 0; i < ndim; ++i) {
    if (dims[i] == 0 || c > kMaxBytes / dims[i]) return bfail(QSMI_ERR_ARG, "grid too large or empty");
    c *= dims[i];
  }
  const size_t total = header_bytes((unsigned)ndim, (unsigned)nVar) + c * sizeof(float);
  if (total > kMaxBytes) return bfail(QSMI_ERR_ARG, "grid above 2^30 bytes");
  *bytes = total;
  return QSMI_OK;
}

int qsmi_bgrid_serialize(int ndim, const unsigned* dims, int nVar, const int* varTypes, int encoding,
                         const float* data, void* out, size_t outBytes) {
  size_t total = 0;
  int rc = qsmi_bgrid_bytes(ndim, dims, nVar, &total);
  if (rc) return rc;
  if (!varTypes || !data || !out) return bfail(QSMI_ERR_ARG, "null argument");
  if (outBytes < total) return bfail(QSMI_ERR_ARG, "output buffer too small");
  if (encoding < 0 || encoding > 2) return bfail(QSMI_ERR_ARG, "encoding must be 0, 1 or 2");
  unsigned char* p = (unsigned char*)out;
  memcpy(p, &total, sizeof(size_t));
  size_t o = sizeof(size_t);
  auto put = [&](uint32_t v) {
    memcpy(p + o, &v, 4);
    o += 4;
  };
  put((uint32_t)ndim);
  for (int i = 0; i < ndim; ++i) put(dims[i]);
  put((uint32_t)nVar);
  for (int v = 0; v < nVar; ++v) {
    if (varTypes[v] != 0 && varTypes[v] != 1) return bfail(QSMI_ERR_ARG, "variable type must be 0 or 1");
    put((uint32_t)varTypes[v]);
  }
  put((uint32_t)encoding);
  memcpy(p + o, data, total - o);
  return QSMI_OK;
}

int qsmi_bgrid_valid(const void* raw, size_t bytes) {
  return parse((const unsigned char*)raw, bytes, nullptr, nullptr, nullptr, nullptr, nullptr) ? 1 : 0;
}

int qsmi_bgrid_header(const void* raw, size_t bytes, int* ndim, unsigned* dims, int* nVar, int* varTypes,
                      int typesCap, int* encoding, size_t* dataOffset) {
  unsigned nd = 0, nv = 0;
  size_t cells = 0, off = 0;
  unsigned d[kMaxDims];
  if (!parse((const unsigned char*)raw, bytes, &nd, d, &nv, &cells, &off))
    return bfail(QSMI_ERR_ARG, "not a valid serialized grid");
  const unsigned char* p = (const unsigned char*)raw;
  if (ndim) *ndim = (int)nd;
  if (dims)
    for (unsigned i = 0; i < nd; ++i) dims[i] = d[i];
  if (nVar) *nVar = (int)nv;
  const size_t tOff = sizeof(size_t) + 4 * (1 + nd + 1);
  if (varTypes) {
    if (typesCap < (int)nv) return bfail(QSMI_ERR_ARG, "variable type array too small");
    for (unsigned v = 0; v < nv; ++v) varTypes[v] = (int)rd32(p + tOff + 4 * v);
  }
  if (encoding) *encoding = (int)rd32(p + tOff + 4 * nv);
  if (dataOffset) *dataOffset = off;
  return QSMI_OK;
}

int qsmi_bgrid_write_file(const char* path, const void* raw, size_t bytes, int gzip) {
  if (!path || !raw) return bfail(QSMI_ERR_ARG, "null argument");
  if (!qsmi_bgrid_valid(raw, bytes)) return bfail(QSMI_ERR_ARG, "not a valid serialized grid");
  if (gzip) {
    gzFile f = gzopen(path, "wb");
    if (!f) return bfail(QSMI_ERR_ARG, std::string("cannot open ") + path);
    size_t done = 0;
    while (done < bytes) {
      const unsigned chunk = (unsigned)std::min<size_t>(bytes - done, size_t(1) << 26);
      if (gzwrite(f, (const char*)raw + done, chunk) != (int)chunk) {
        gzclose(f);
        return bfail(QSMI_ERR_ARG, std::string("write failed: ") + path);
      }
      done += chunk;
    }
    if (gzclose(f) != Z_OK) return bfail(QSMI_ERR_ARG, std::string("write failed: ") + path);
    return QSMI_OK;
  }
  FILE* f = fopen(path, "wb");
  if (!f) return bfail(QSMI_ERR_ARG, std::string("cannot open ") + path);
  const bool ok = fwrite(raw, 1, bytes, f) == bytes;
  if (fclose(f) != 0 || !ok) return bfail(QSMI_ERR_ARG, std::string("write failed: ") + path);
  return QSMI_OK;
}

int qsmi_bgrid_read_file(const char* path, void* out, size_t outBytes, size_t* bytes) {
  if (!path || !bytes) return bfail(QSMI_ERR_ARG, "null argument");
  std::vector<unsigned char> buf;
  // gzopen reads plain files transparently, so one reader serves both forms
  if (!read_all_gz(path, buf)) return bfail(QSMI_ERR_ARG, std::string("cannot read ") + path);
  if (!qsmi_bgrid_valid(buf.data(), buf.size())) return bfail(QSMI_ERR_ARG, std::string("invalid grid file ") + path);
  *bytes = buf.size();
  if (!out) return QSMI_OK;  // size query
  if (outBytes < buf.size()) return bfail(QSMI_ERR_ARG, "output buffer too small");
  memcpy(out, buf.data(), buf.size());
  return QSMI_OK;
}

}  // extern "C"

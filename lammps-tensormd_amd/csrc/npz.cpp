#include "npz.h"

#include <zlib.h>

#include <cstdint>
#include <cstring>
#include <fstream>

namespace tmd {

namespace {

uint16_t rd16(const unsigned char *p) { return (uint16_t) (p[0] | (p[1] << 8)); }
uint32_t rd32(const unsigned char *p)
{
  return (uint32_t) p[0] | ((uint32_t) p[1] << 8) | ((uint32_t) p[2] << 16) |
      ((uint32_t) p[3] << 24);
}
uint64_t rd64(const unsigned char *p) { return (uint64_t) rd32(p) | ((uint64_t) rd32(p + 4) << 32); }

bool parse_npy(const unsigned char *p, size_t n, NpyArray &a, std::string &err)
{
  static const unsigned char magic[6] = {0x93, 'N', 'U', 'M', 'P', 'Y'};
  if (n < 10 || std::memcmp(p, magic, 6) != 0) {
    err = "not an npy member";
    return false;
  }
  int major = p[6];
  size_t hlen, hoff;
  if (major == 1) {
    hlen = rd16(p + 8);
    hoff = 10;
  } else {
    if (n < 12) {
      err = "truncated npy header";
      return false;
    }
    hlen = rd32(p + 8);
    hoff = 12;
  }
  if (hoff + hlen > n) {
    err = "truncated npy header";
    return false;
  }
  std::string h((const char *) p + hoff, hlen);
  auto field = [&](const char *key) -> std::string {
    size_t k = h.find(key);
    if (k == std::string::npos) return "";
    k = h.find(':', k);
    if (k == std::string::npos) return "";
    return h.substr(k + 1);
  };
  std::string d = field("'descr'");
  size_t q1 = d.find('\'');
  size_t q2 = d.find('\'', q1 + 1);
  if (q1 == std::string::npos || q2 == std::string::npos) {
    err = "bad npy descr";
    return false;
  }
  std::string descr = d.substr(q1 + 1, q2 - q1 - 1);
  if (descr.size() < 3 || descr[0] == '>') {
    err = "unsupported npy dtype " + descr;
    return false;
  }
  a.kind = descr[1];
  a.word = std::atoi(descr.c_str() + 2);
  if (!(a.kind == 'f' || a.kind == 'i' || a.kind == 'u' || a.kind == 'b') || a.word <= 0) {
    err = "unsupported npy dtype " + descr;
    return false;
  }
  std::string fo = field("'fortran_order'");
  a.fortran = fo.find("True") != std::string::npos && fo.find("True") < fo.find(',');
  std::string sh = field("'shape'");
  size_t l = sh.find('('), r = sh.find(')');
  if (l == std::string::npos || r == std::string::npos) {
    err = "bad npy shape";
    return false;
  }
  a.shape.clear();
  std::string inner = sh.substr(l + 1, r - l - 1);
  size_t pos = 0;
  while (pos < inner.size()) {
    while (pos < inner.size() && (inner[pos] == ' ' || inner[pos] == ',')) pos++;
    if (pos >= inner.size()) break;
    a.shape.push_back((size_t) std::strtoull(inner.c_str() + pos, nullptr, 10));
    while (pos < inner.size() && inner[pos] != ',') pos++;
  }
  size_t nbytes = a.size() * (size_t) a.word;
  size_t doff = hoff + hlen;
  if (doff + nbytes > n) {
    err = "truncated npy data";
    return false;
  }
  if (a.fortran && a.shape.size() > 1) {
    err = "fortran-order arrays are not supported";
    return false;
  }
  a.bytes.assign(p + doff, p + doff + nbytes);
  return true;
}

}  // namespace

double NpyArray::get_double(size_t i) const
{
  const unsigned char *p = bytes.data() + i * word;
  if (kind == 'f') {
    if (word == 8) {
      double v;
      std::memcpy(&v, p, 8);
      return v;
    }
    if (word == 4) {
      float v;
      std::memcpy(&v, p, 4);
      return v;
    }
  }
  return (double) get_int(i);
}

long long NpyArray::get_int(size_t i) const
{
  const unsigned char *p = bytes.data() + i * word;
  if (kind == 'f') return (long long) get_double(i);
  long long v = 0;
  if (kind == 'i') {
    if (word == 1) v = (int8_t) p[0];
    if (word == 2) v = (int16_t) rd16(p);
    if (word == 4) v = (int32_t) rd32(p);
    if (word == 8) v = (int64_t) rd64(p);
  } else {
    if (word == 1) v = p[0];
    if (word == 2) v = rd16(p);
    if (word == 4) v = rd32(p);
    if (word == 8) v = (long long) rd64(p);
  }
  return v;
}

std::vector<double> NpyArray::as_doubles() const
{
  std::vector<double> v(size());
  for (size_t i = 0; i < v.size(); i++) v[i] = get_double(i);
  return v;
}

std::vector<int> NpyArray::as_ints() const
{
  std::vector<int> v(size());
  for (size_t i = 0; i < v.size(); i++) v[i] = (int) get_int(i);
  return v;
}

bool read_npz(const std::string &path, Npz &out, std::string &err)
{
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    err = "cannot open " + path;
    return false;
  }
  std::vector<unsigned char> buf((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  size_t n = buf.size();
  if (n < 22) {
    err = "not a zip file: " + path;
    return false;
  }
  // end of central directory record
  long eocd = -1;
  for (long i = (long) n - 22; i >= 0 && i >= (long) n - 22 - 65536; i--)
    if (rd32(&buf[i]) == 0x06054b50) {
      eocd = i;
      break;
    }
  if (eocd < 0) {
    err = "zip end record not found: " + path;
    return false;
  }
  uint64_t nent = rd16(&buf[eocd + 10]);
  uint64_t cdoff = rd32(&buf[eocd + 16]);
  if (cdoff == 0xFFFFFFFFu || nent == 0xFFFF) {
    // zip64 end of central directory locator precedes the EOCD
    if (eocd >= 20 && rd32(&buf[eocd - 20]) == 0x07064b50) {
      uint64_t z64 = rd64(&buf[eocd - 20 + 8]);
      if (z64 + 56 <= n && rd32(&buf[z64]) == 0x06064b50) {
        nent = rd64(&buf[z64 + 32]);
        cdoff = rd64(&buf[z64 + 48]);
      }
    }
  }
  size_t p = (size_t) cdoff;
  for (uint64_t e = 0; e < nent; e++) {
    if (p + 46 > n || rd32(&buf[p]) != 0x02014b50) {
      err = "corrupt zip central directory: " + path;
      return false;
    }
    int method = rd16(&buf[p + 10]);
    uint64_t csize = rd32(&buf[p + 20]);
    uint64_t usize = rd32(&buf[p + 24]);
    int nlen = rd16(&buf[p + 28]), xlen = rd16(&buf[p + 30]), clen = rd16(&buf[p + 32]);
    uint64_t loff = rd32(&buf[p + 42]);
    std::string name((const char *) &buf[p + 46], nlen);
    // zip64 extra: fields present in order usize, csize, loff for those = 0xFFFFFFFF
    size_t x = p + 46 + nlen, xe = x + xlen;
    while (x + 4 <= xe) {
      int id = rd16(&buf[x]), sz = rd16(&buf[x + 2]);
      if (id == 0x0001) {
        size_t q = x + 4;
        if (usize == 0xFFFFFFFFu && q + 8 <= x + 4 + sz) usize = rd64(&buf[q]), q += 8;
        if (csize == 0xFFFFFFFFu && q + 8 <= x + 4 + sz) csize = rd64(&buf[q]), q += 8;
        if (loff == 0xFFFFFFFFu && q + 8 <= x + 4 + sz) loff = rd64(&buf[q]), q += 8;
      }
      x += 4 + sz;
    }
    p = xe + clen;
    if (loff + 30 > n || rd32(&buf[loff]) != 0x04034b50) {
      err = "corrupt zip local header: " + path;
      return false;
    }
    size_t data = loff + 30 + rd16(&buf[loff + 26]) + rd16(&buf[loff + 28]);
    if (data + csize > n) {
      err = "truncated zip member " + name;
      return false;
    }
    std::vector<unsigned char> raw;
    const unsigned char *src;
    size_t srcn;
    if (method == 0) {
      src = &buf[data];
      srcn = csize;
    } else if (method == 8) {
      raw.resize(usize);
      z_stream zs;
      std::memset(&zs, 0, sizeof(zs));
      if (inflateInit2(&zs, -15) != Z_OK) {
        err = "zlib init failed";
        return false;
      }
      zs.next_in = &buf[data];
      zs.avail_in = (uInt) csize;
      zs.next_out = raw.data();
      zs.avail_out = (uInt) usize;
      int rc = inflate(&zs, Z_FINISH);
      inflateEnd(&zs);
      if (rc != Z_STREAM_END) {
        err = "inflate failed for " + name;
        return false;
      }
      src = raw.data();
      srcn = usize;
    } else {
      err = "unsupported zip compression for " + name;
      return false;
    }
    if (name.size() > 4 && name.substr(name.size() - 4) == ".npy") name = name.substr(0, name.size() - 4);
    NpyArray a;
    std::string e2;
    if (!parse_npy(src, srcn, a, e2)) {
      err = name + ": " + e2;
      return false;
    }
    out[name] = std::move(a);
  }
  return true;
}

}  // namespace tmd

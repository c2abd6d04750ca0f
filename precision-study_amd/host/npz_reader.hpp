// npz_reader.hpp -- reads the float32 arrays of an uncompressed .npz file
// (numpy.savez: a zip archive of .npy members, each stored without
// compression), as the learned selector's model file is written.
//
// Only what that format needs: zip local file headers with method 0
// (stored), .npy version 1 or 2 headers, little-endian float32 ('<f4')
// arrays in C order.  Members of any other type are skipped; a malformed
// or compressed archive is an error.
#ifndef MHD_NPZ_READER_HPP
#define MHD_NPZ_READER_HPP

#include <cstdint>
#include <cstring>
#include <fstream>
#include <iterator>
#include <map>
#include <string>
#include <vector>

namespace mhd {

struct NpzArray {
  std::vector<size_t> shape;  // empty: a scalar
  std::vector<float> data;
};

namespace npz_detail {

inline uint32_t le32(const unsigned char* p) {
  return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24;
}
inline uint16_t le16(const unsigned char* p) { return (uint16_t)(p[0] | p[1] << 8); }

// Parses one .npy member; returns false if it is malformed, and sets
// `wanted` to whether it holds a little-endian float32 C-order array.
inline bool parse_npy(const unsigned char* p, size_t n, NpzArray& out, bool& wanted,
                      std::string& err) {
  static const unsigned char kMagic[6] = {0x93, 'N', 'U', 'M', 'P', 'Y'};
  if (n < 10 || std::memcmp(p, kMagic, 6) != 0) {
    err = "member is not a .npy array";
    return false;
  }
  const int major = p[6];
  size_t hlen, off;
  if (major == 1) {
    hlen = le16(p + 8);
    off = 10;
  } else if (major == 2 || major == 3) {
    if (n < 12) {
      err = "truncated .npy header";
      return false;
    }
    hlen = le32(p + 8);
    off = 12;
  } else {
    err = "unknown .npy version";
    return false;
  }
  if (off + hlen > n) {
    err = "truncated .npy header";
    return false;
  }
  const std::string h((const char*)p + off, hlen);
  wanted = h.find("'descr': '<f4'") != std::string::npos &&
           h.find("'fortran_order': False") != std::string::npos;
  if (!wanted) return true;
  const size_t s0 = h.find("'shape': (");
  if (s0 == std::string::npos) {
    err = "no shape in .npy header";
    return false;
  }
  const size_t s1 = h.find(')', s0);
  if (s1 == std::string::npos) {
    err = "bad shape in .npy header";
    return false;
  }
  out.shape.clear();
  size_t count = 1;
  const std::string dims = h.substr(s0 + 10, s1 - (s0 + 10));
  size_t q = 0;
  while (q < dims.size()) {
    while (q < dims.size() && (dims[q] == ' ' || dims[q] == ',')) q++;
    if (q >= dims.size()) break;
    size_t v = 0;
    bool digit = false;
    while (q < dims.size() && dims[q] >= '0' && dims[q] <= '9') {
      v = v * 10 + (size_t)(dims[q] - '0');
      q++;
      digit = true;
      if (v > n) {  // no dimension exceeds the member's byte count
        err = "bad shape in .npy header";
        return false;
      }
    }
    if (!digit) {
      err = "bad shape in .npy header";
      return false;
    }
    out.shape.push_back(v);
    // A shape whose byte count does not fit the member is malformed; the
    // check before the product keeps it from wrapping.
    if (v != 0 && count > n / 4 / v) {
      err = "truncated .npy data";
      return false;
    }
    count *= v;
  }
  const size_t data_off = off + hlen;
  if (data_off + count * 4 > n) {
    err = "truncated .npy data";
    return false;
  }
  out.data.resize(count);
  for (size_t k = 0; k < count; k++) {
    const uint32_t bits = le32(p + data_off + 4 * k);
    std::memcpy(&out.data[k], &bits, 4);
  }
  return true;
}

}  // namespace npz_detail

// Reads every float32 member of `path` into `out` (keyed by the member
// name without ".npy").  Returns false with a message in `err` on an
// unreadable, compressed or malformed archive.
inline bool read_npz_f32(const std::string& path, std::map<std::string, NpzArray>& out,
                         std::string& err) {
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    err = "cannot open " + path;
    return false;
  }
  const std::vector<unsigned char> b((std::istreambuf_iterator<char>(f)),
                                     std::istreambuf_iterator<char>());
  size_t p = 0;
  int members = 0;
  while (p + 30 <= b.size() && npz_detail::le32(&b[p]) == 0x04034b50u) {
    const unsigned char* h = &b[p];
    const uint16_t flags = npz_detail::le16(h + 6), method = npz_detail::le16(h + 8);
    uint64_t csize = npz_detail::le32(h + 18), usize = npz_detail::le32(h + 22);
    const uint16_t nlen = npz_detail::le16(h + 26), xlen = npz_detail::le16(h + 28);
    if (p + 30 + (size_t)nlen + xlen > b.size()) {
      err = "truncated zip header";
      return false;
    }
    const std::string name((const char*)h + 30, nlen);
    // ZIP64 sizes (numpy writes them for large members) sit in the extra
    // field, id 0x0001: uncompressed then compressed size.
    if (csize == 0xFFFFFFFFu || usize == 0xFFFFFFFFu) {
      size_t x = p + 30 + nlen;
      const size_t xend = x + xlen;
      bool found = false;
      while (x + 4 <= xend) {
        const uint16_t id = npz_detail::le16(&b[x]), len = npz_detail::le16(&b[x + 2]);
        if (id == 0x0001 && len >= 16 && x + 4 + 16 <= xend) {
          usize = npz_detail::le32(&b[x + 4]) | (uint64_t)npz_detail::le32(&b[x + 8]) << 32;
          csize = npz_detail::le32(&b[x + 12]) | (uint64_t)npz_detail::le32(&b[x + 16]) << 32;
          found = true;
          break;
        }
        x += 4 + (size_t)len;
      }
      if (!found) {
        err = "zip64 member without sizes: " + name;
        return false;
      }
    }
    if (method != 0 || (flags & 0x8) != 0 || csize != usize) {
      err = "compressed or streamed zip member (write the model with numpy.savez): " + name;
      return false;
    }
    const size_t data = p + 30 + nlen + xlen;
    if (data + csize > b.size()) {
      err = "truncated zip member: " + name;
      return false;
    }
    NpzArray arr;
    bool wanted = false;
    if (!npz_detail::parse_npy(&b[data], (size_t)csize, arr, wanted, err)) {
      err = name + ": " + err;
      return false;
    }
    if (wanted) {
      std::string key = name;
      if (key.size() > 4 && key.compare(key.size() - 4, 4, ".npy") == 0)
        key.resize(key.size() - 4);
      out[key] = arr;
    }
    members++;
    p = data + (size_t)csize;
  }
  if (members == 0) {
    err = "not a zip archive: " + path;
    return false;
  }
  return true;
}

}  // namespace mhd

#endif  // MHD_NPZ_READER_HPP

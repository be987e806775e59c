// wg_pq_stage.cpp — Parquet column chunks staged for the device decode, on
// host threads. Host code only: built by the host compiler.
//
// The reference's ColumnPageReader reads a chunk page by page: it parses the
// Thrift compact page header (DecodePageHeader, internal/storage/parquet/
// thrift.go:1156-1330), decompresses the body (Decompress, decompress.go:
// 15-110) and decodes levels and values (decodeDataPageV1 / V2,
// page_reader.go:434-560). The value decode runs on the device here
// (wg_parquet.hip); this file does the rest for whole column chunks:
//   walk    per chunk, the page headers alone (a small read at each header):
//           every page's type and sizes, and where its streams will sit in
//           the chunk's staged region; the regions are laid end to end;
//   stage   per chunk, on a pool of threads: one read of the chunk, each page
//           body decompressed straight to its planned place in a pinned slot,
//           the run headers of each hybrid stream scanned, and the slot
//           copied to the chunk's region of one device buffer asynchronously
//           while the thread goes on to its next chunk (a slot per thread,
//           written again once its copy has completed);
//   tables  the run tables of all chunks in one copy, then the page table
//           with device addresses.
// The device reads streams as 64-bit words: each starts on an 8-byte boundary
// and has 16 readable bytes after it. A v1 page is decompressed as one piece
// (length-prefixed levels, then the values); its levels and dictionary
// indices are read in place, from the 8-byte word at or below their first
// byte, with the bit offsets of their bit-packed runs raised by the distance;
// PLAIN values that follow levels are moved up to the next 8-byte boundary.
//
// Decompressors: snappy and LZ4 (block and frame formats) are written here
// from their published format descriptions; gzip is zlib's inflate; zstd is
// the system libzstd, loaded on first use.
#include <dlfcn.h>
#include <fcntl.h>
#include <sched.h>
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "wg_host.hpp"

namespace wg {
namespace {

// ----------------------------------------------------------------- errors

// First failure of a pool of threads; reported on the calling thread (the
// library's error message is per thread).
struct Failure {
  std::mutex mu;
  std::atomic<bool> failed{false};
  wg_status st = WG_OK;
  std::string msg;

  wg_status set(wg_status s, const char* fmt, ...) __attribute__((format(printf, 3, 4))) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    std::lock_guard<std::mutex> g(mu);
    if (!failed.load()) {
      st = s;
      msg = buf;
      failed.store(true);
    }
    return s;
  }
  wg_status report() {
    if (failed.load()) set_error("%s", msg.c_str());
    return failed.load() ? st : WG_OK;
  }
};

#define PQ_FAIL(F, code, ...) return (F).set((code), __VA_ARGS__)

inline uint64_t align8(uint64_t v) { return (v + 7) & ~7ull; }

// --------------------------------------------------------- decompression

constexpr int kCodecNone = 0, kCodecSnappy = 1, kCodecGzip = 2, kCodecLz4 = 5, kCodecZstd = 6, kCodecLz4Raw = 7;

// Back-reference copy of `len` bytes from `dist` bytes behind op; the regions
// overlap when dist < len (a run repeating its last `dist` bytes).
inline void copy_back(uint8_t* op, uint64_t dist, uint64_t len, const uint8_t* end) {
  const uint8_t* src = op - dist;
  if (dist >= 8) {
    uint64_t i = 0;
    for (; i + 8 <= len && op + i + 8 <= end; i += 8) memcpy(op + i, src + i, 8);
    for (; i < len; ++i) op[i] = src[i];
  } else {
    for (uint64_t i = 0; i < len; ++i) op[i] = src[i];
  }
}

// Snappy block format (format_description.txt of the snappy project): a
// varint of the uncompressed length, then elements whose tag's low two bits
// select a literal (length in the tag or in 1-4 following bytes) or a copy
// with a 1-, 2- or 4-byte offset.
bool snappy_decode(const uint8_t* ip, uint64_t n, uint8_t* dst, uint64_t dst_len, std::string* why) {
  const uint8_t* const iend = ip + n;
  uint64_t want = 0;
  for (int s = 0;; s += 7) {
    if (ip >= iend || s > 35) return *why = "snappy length varint", false;
    const uint8_t b = *ip++;
    want |= (uint64_t)(b & 0x7F) << s;
    if (!(b & 0x80)) break;
  }
  if (want != dst_len) return *why = "snappy length " + std::to_string(want) + " != page size", false;
  uint8_t* op = dst;
  uint8_t* const oend = dst + dst_len;
  while (ip < iend) {
    const uint8_t tag = *ip++;
    uint64_t len, dist;
    switch (tag & 3) {
      case 0: {
        len = (tag >> 2) + 1;
        if (len > 60) {
          const int k = (int)len - 60;  // 1..4 little-endian length bytes
          if (iend - ip < k) return *why = "snappy literal length", false;
          len = 0;
          for (int i = 0; i < k; ++i) len |= (uint64_t)ip[i] << (8 * i);
          len += 1;
          ip += k;
        }
        if ((uint64_t)(iend - ip) < len || (uint64_t)(oend - op) < len) return *why = "snappy literal", false;
        memcpy(op, ip, len);
        ip += len;
        op += len;
        continue;
      }
      case 1:
        if (ip >= iend) return *why = "snappy copy", false;
        len = 4 + ((tag >> 2) & 7);
        dist = ((uint64_t)(tag >> 5) << 8) | *ip++;
        break;
      case 2:
        if (iend - ip < 2) return *why = "snappy copy", false;
        len = 1 + (tag >> 2);
        dist = (uint64_t)ip[0] | (uint64_t)ip[1] << 8;
        ip += 2;
        break;
      default:
        if (iend - ip < 4) return *why = "snappy copy", false;
        len = 1 + (tag >> 2);
        dist = (uint64_t)ip[0] | (uint64_t)ip[1] << 8 | (uint64_t)ip[2] << 16 | (uint64_t)ip[3] << 24;
        ip += 4;
        break;
    }
    if (dist == 0 || dist > (uint64_t)(op - dst) || (uint64_t)(oend - op) < len)
      return *why = "snappy copy out of range", false;
    copy_back(op, dist, len, oend);
    op += len;
  }
  if (op != oend) return *why = "snappy output short of the page size", false;
  return true;
}

// One LZ4 block (lz4_Block_format.md of the LZ4 project): sequences of a
// token (literal length, match length - 4, each 15 = continued in bytes of
// 255), the literals, a 2-byte offset and the match; the last sequence ends
// after its literals. `base` is where back references may reach (the start of
// the frame's output for linked blocks). Returns the bytes written.
bool lz4_block(const uint8_t* ip, uint64_t n, const uint8_t* base, uint8_t* op, uint8_t* oend, uint64_t* wrote,
               std::string* why) {
  const uint8_t* const iend = ip + n;
  uint8_t* const o0 = op;
  while (ip < iend) {
    const uint8_t token = *ip++;
    uint64_t lit = token >> 4;
    if (lit == 15) {
      uint8_t b;
      do {
        if (ip >= iend) return *why = "lz4 literal length", false;
        b = *ip++;
        lit += b;
      } while (b == 255);
    }
    if ((uint64_t)(iend - ip) < lit || (uint64_t)(oend - op) < lit) return *why = "lz4 literals", false;
    memcpy(op, ip, lit);
    ip += lit;
    op += lit;
    if (ip == iend) break;  // the last sequence: literals only
    if (iend - ip < 2) return *why = "lz4 offset", false;
    const uint64_t dist = (uint64_t)ip[0] | (uint64_t)ip[1] << 8;
    ip += 2;
    uint64_t mlen = (token & 15);
    if (mlen == 15) {
      uint8_t b;
      do {
        if (ip >= iend) return *why = "lz4 match length", false;
        b = *ip++;
        mlen += b;
      } while (b == 255);
    }
    mlen += 4;
    if (dist == 0 || dist > (uint64_t)(op - base) || (uint64_t)(oend - op) < mlen)
      return *why = "lz4 match out of range", false;
    copy_back(op, dist, mlen, oend);
    op += mlen;
  }
  *wrote = (uint64_t)(op - o0);
  return true;
}

// LZ4 frame format (lz4_Frame_format.md): magic, frame descriptor (flags,
// block maximum, optional content size and dictionary id, header checksum),
// blocks of a 4-byte size (high bit = stored uncompressed, 0 = end) with
// optional block checksums, optional content checksum; frames may follow one
// another. Checksums are skipped, not verified.
bool lz4_frame_decode(const uint8_t* ip, uint64_t n, uint8_t* dst, uint64_t dst_len, std::string* why) {
  const uint8_t* const iend = ip + n;
  uint8_t* op = dst;
  uint8_t* const oend = dst + dst_len;
  auto le32 = [](const uint8_t* p) {
    return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24;
  };
  while (ip < iend) {
    if (iend - ip < 7) return *why = "lz4 frame header", false;
    const uint32_t magic = le32(ip);
    ip += 4;
    if ((magic & 0xFFFFFFF0u) == 0x184D2A50u) {  // skippable frame
      const uint32_t sz = le32(ip);
      ip += 4;
      if ((uint64_t)(iend - ip) < sz) return *why = "lz4 skippable frame", false;
      ip += sz;
      continue;
    }
    if (magic != 0x184D2204u) return *why = "lz4 frame magic", false;
    const uint8_t flg = ip[0];
    ip += 2;  // FLG, BD
    if ((flg >> 6) != 1) return *why = "lz4 frame version", false;
    const bool block_sum = flg & 0x10, content_sum = flg & 0x04;
    const uint64_t extra = ((flg & 0x08) ? 8 : 0) + ((flg & 0x01) ? 4 : 0) + 1;
    if ((uint64_t)(iend - ip) < extra) return *why = "lz4 frame descriptor", false;
    ip += extra;
    uint8_t* const frame0 = op;
    for (;;) {
      if (iend - ip < 4) return *why = "lz4 block size", false;
      const uint32_t bs = le32(ip);
      ip += 4;
      if (bs == 0) break;
      const uint64_t sz = bs & 0x7FFFFFFFu;
      if ((uint64_t)(iend - ip) < sz) return *why = "lz4 block", false;
      if (bs & 0x80000000u) {
        if ((uint64_t)(oend - op) < sz) return *why = "lz4 stored block", false;
        memcpy(op, ip, sz);
        op += sz;
      } else {
        uint64_t w = 0;
        if (!lz4_block(ip, sz, frame0, op, oend, &w, why)) return false;
        op += w;
      }
      ip += sz;
      if (block_sum) {
        if (iend - ip < 4) return *why = "lz4 block checksum", false;
        ip += 4;
      }
    }
    if (content_sum) {
      if (iend - ip < 4) return *why = "lz4 content checksum", false;
      ip += 4;
    }
  }
  if (op != oend) return *why = "lz4 output short of the page size", false;
  return true;
}

// The deprecated LZ4 codec: an LZ4 frame (magic 0x184D2204, as the
// reference's reader takes it), else the Hadoop framing parquet-mr and Arrow
// write (per block: big-endian uncompressed and compressed sizes, then a raw
// block), else one raw block.
bool lz4_decode(const uint8_t* ip, uint64_t n, uint8_t* dst, uint64_t dst_len, std::string* why) {
  if (n >= 4 && ip[0] == 0x04 && ip[1] == 0x22 && ip[2] == 0x4D && ip[3] == 0x18)
    return lz4_frame_decode(ip, n, dst, dst_len, why);
  auto be32 = [](const uint8_t* p) {
    return (uint64_t)p[0] << 24 | (uint64_t)p[1] << 16 | (uint64_t)p[2] << 8 | (uint64_t)p[3];
  };
  const uint8_t* p = ip;
  const uint8_t* const iend = ip + n;
  uint8_t* op = dst;
  bool hadoop = n >= 8;
  while (hadoop && p < iend) {
    if (iend - p < 8) {
      hadoop = false;
      break;
    }
    const uint64_t usz = be32(p), csz = be32(p + 4);
    if (csz > (uint64_t)(iend - p - 8) || usz > (uint64_t)(dst + dst_len - op)) {
      hadoop = false;
      break;
    }
    uint64_t w = 0;
    std::string ignored;
    if (!lz4_block(p + 8, csz, op, op, op + usz, &w, &ignored) || w != usz) {
      hadoop = false;
      break;
    }
    op += usz;
    p += 8 + csz;
  }
  if (hadoop && op == dst + dst_len) return true;
  uint64_t w = 0;
  if (!lz4_block(ip, n, dst, dst, dst + dst_len, &w, why)) return false;
  if (w != dst_len) return *why = "lz4 output short of the page size", false;
  return true;
}

// gzip (RFC 1952) members through zlib, one after another until the page is full.
bool gzip_decode(const uint8_t* ip, uint64_t n, uint8_t* dst, uint64_t dst_len, std::string* why) {
  if (n > 0xFFFFFFFFull || dst_len > 0xFFFFFFFFull) return *why = "gzip page over 4 GiB", false;
  z_stream z;
  memset(&z, 0, sizeof(z));
  if (inflateInit2(&z, 15 + 32) != Z_OK) return *why = "inflateInit2", false;  // gzip (or zlib) header
  z.next_in = const_cast<Bytef*>(ip);
  z.avail_in = (uInt)n;
  z.next_out = dst;
  z.avail_out = (uInt)dst_len;
  bool ok = false;
  for (;;) {
    const uInt in0 = z.avail_in, out0 = z.avail_out;
    const int r = inflate(&z, Z_FINISH);
    if (r == Z_STREAM_END) {
      if (z.avail_out == 0 || z.avail_in == 0) {
        ok = z.avail_out == 0;
        break;
      }
      if (inflateReset(&z) != Z_OK) break;  // another member follows
      continue;
    }
    if ((r == Z_BUF_ERROR || r == Z_OK) && z.avail_out == 0) {
      ok = true;  // full page; trailing bytes ignored as a limited reader would
      break;
    }
    if (r == Z_OK && (z.avail_in != in0 || z.avail_out != out0)) continue;
    *why = std::string("gzip: ") + (z.msg ? z.msg : "corrupt or truncated stream");
    inflateEnd(&z);
    return false;
  }
  inflateEnd(&z);
  if (!ok) *why = "gzip output short of the page size";
  return ok;
}

// zstd from the system library, resolved once.
struct Zstd {
  size_t (*decompress)(void*, size_t, const void*, size_t) = nullptr;
  unsigned (*is_error)(size_t) = nullptr;
  const char* (*error_name)(size_t) = nullptr;
};
const Zstd* zstd() {
  static Zstd z;
  static std::once_flag once;
  std::call_once(once, [] {
    void* h = dlopen("libzstd.so.1", RTLD_NOW | RTLD_LOCAL);
    if (!h) h = dlopen("libzstd.so", RTLD_NOW | RTLD_LOCAL);
    if (!h) return;
    z.decompress = (size_t(*)(void*, size_t, const void*, size_t))dlsym(h, "ZSTD_decompress");
    z.is_error = (unsigned (*)(size_t))dlsym(h, "ZSTD_isError");
    z.error_name = (const char* (*)(size_t))dlsym(h, "ZSTD_getErrorName");
  });
  return z.decompress && z.is_error ? &z : nullptr;
}

// Page body `ip[0..n)` of codec `codec` into exactly dst_len bytes.
wg_status decompress(int codec, const uint8_t* ip, uint64_t n, uint8_t* dst, uint64_t dst_len, std::string* why) {
  bool ok = false;
  switch (codec) {
    case kCodecNone:
      if (n < dst_len) return *why = "stored page shorter than its size", WG_ERR_INVALID;
      if (dst_len) memcpy(dst, ip, dst_len);
      return WG_OK;
    case kCodecSnappy: ok = snappy_decode(ip, n, dst, dst_len, why); break;
    case kCodecGzip: ok = gzip_decode(ip, n, dst, dst_len, why); break;
    case kCodecLz4: ok = lz4_decode(ip, n, dst, dst_len, why); break;
    case kCodecLz4Raw: {
      uint64_t w = 0;
      ok = lz4_block(ip, n, dst, dst, dst + dst_len, &w, why) && w == dst_len;
      if (ok || !why->empty()) break;
      *why = "lz4 output short of the page size";
      break;
    }
    case kCodecZstd: {
      const Zstd* z = zstd();
      if (!z) return *why = "zstd: libzstd.so.1 not found", WG_ERR_UNSUPPORTED;
      const size_t r = z->decompress(dst, dst_len, ip, n);
      if (z->is_error(r)) {
        *why = std::string("zstd: ") + (z->error_name ? z->error_name(r) : "corrupt frame");
        return WG_ERR_INVALID;
      }
      ok = r == dst_len;
      if (!ok) *why = "zstd output short of the page size";
      break;
    }
    default:
      *why = "compression codec " + std::to_string(codec) + " not on the GPU decode path";
      return WG_ERR_UNSUPPORTED;
  }
  return ok ? WG_OK : WG_ERR_INVALID;
}

// ------------------------------------------------------------- hybrid runs

// Run headers of one RLE / bit-packing hybrid stream (rle.go:50-120): a
// varint header; header & 1 = 0: an RLE run of header >> 1 values, the value
// in ceil(bw / 8) bytes; else a bit-packed run of (header >> 1) groups of 8
// values, bw bytes per group. Appends {first value, count, RLE value or bit
// offset + bias, packed} per run; *ones = values equal to 1 among the first
// `count` (the present rows of a definition-level stream).
wg_status scan_runs(const uint8_t* data, uint64_t len, int bw, uint64_t count, uint32_t bias, std::vector<uint32_t>* runs,
                    uint64_t* used, uint64_t* ones, std::string* why) {
  if (bw < 0 || bw > 32) return *why = "bit width " + std::to_string(bw), WG_ERR_INVALID;
  if (count >= (1ull << 32) || len >= (1ull << 29)) return *why = "hybrid stream too long", WG_ERR_UNSUPPORTED;
  const int vbytes = (bw + 7) / 8;
  uint64_t off = 0, got = 0, n1 = 0;
  while (got < count) {
    if (off >= len) {
      *why = "hybrid stream ends after " + std::to_string(got) + " of " + std::to_string(count) + " values";
      return WG_ERR_INVALID;
    }
    uint64_t header = 0;
    for (int s = 0;; s += 7) {
      if (off >= len || s >= 64) return *why = "bad run header varint", WG_ERR_INVALID;
      const uint8_t b = data[off++];
      header |= (uint64_t)(b & 0x7F) << s;
      if (!(b & 0x80)) break;
    }
    uint64_t n;
    uint32_t payload = 0, packed = 0;
    if (header & 1) {
      const uint64_t groups = header >> 1;
      n = groups * 8;
      payload = (uint32_t)(off * 8) + bias;
      packed = 1;
      if (off + groups * (uint64_t)bw > len) return *why = "bit-packed run past the stream", WG_ERR_INVALID;
      if (bw == 1 && ones) {  // ones among this run's values below `count`
        const uint64_t take = std::min<uint64_t>(n, count - got);
        const uint8_t* p = data + off;
        uint64_t full = take / 8, i = 0;
        for (; i + 8 <= full; i += 8) {
          uint64_t w;
          memcpy(&w, p + i, 8);
          n1 += (uint64_t)__builtin_popcountll(w);
        }
        for (; i < full; ++i) n1 += (uint64_t)__builtin_popcount(p[i]);
        if (take & 7) n1 += (uint64_t)__builtin_popcount(p[full] & ((1u << (take & 7)) - 1));
      }
      off += groups * (uint64_t)bw;
    } else {
      n = header >> 1;
      if (off + vbytes > len) return *why = "RLE value past the stream", WG_ERR_INVALID;
      for (int i = 0; i < vbytes; ++i) payload |= (uint32_t)data[off + i] << (8 * i);
      off += vbytes;
      if (payload == 1) n1 += std::min<uint64_t>(n, count - got);
    }
    if (n == 0) continue;
    if (runs) {
      runs->push_back((uint32_t)got);
      runs->push_back((uint32_t)std::min<uint64_t>(n, count - got));
      runs->push_back(payload);
      runs->push_back(packed);
    }
    got += n;  // a bit-packed run may pad its last group past `count`
  }
  *used = off;
  if (ones) *ones = n1;
  return WG_OK;
}

// ------------------------------------------------------- Thrift page header

enum : int { kCtStop = 0, kCtTrue = 1, kCtFalse = 2, kCtByte = 3, kCtI16 = 4, kCtI32 = 5, kCtI64 = 6, kCtDouble = 7,
             kCtBinary = 8, kCtList = 9, kCtSet = 10, kCtMap = 11, kCtStruct = 12 };

// Thrift compact protocol reader (the encoding thrift.go's decoder reads):
// field headers carry a field-id delta and a type; integers are zigzag
// varints; unknown fields are skipped by type.
struct Compact {
  const uint8_t* p;
  const uint8_t* end;
  bool bad = false;

  uint8_t byte() {
    if (p >= end) {
      bad = true;
      return 0;
    }
    return *p++;
  }
  uint64_t varint() {
    uint64_t v = 0;
    for (int s = 0; s < 70; s += 7) {
      const uint8_t b = byte();
      if (bad) return 0;
      v |= (uint64_t)(b & 0x7F) << s;
      if (!(b & 0x80)) return v;
    }
    bad = true;
    return 0;
  }
  int64_t zigzag() {
    const uint64_t v = varint();
    return (int64_t)(v >> 1) ^ -(int64_t)(v & 1);
  }
  void skip_bytes(uint64_t n) {
    if ((uint64_t)(end - p) < n) {
      bad = true;
      p = end;
    } else {
      p += n;
    }
  }
  // skip one value of type t (`elem` = inside a list/set/map, where a bool is a byte)
  void skip(int t, int depth, bool elem = false) {
    if (depth > 32) {
      bad = true;
      return;
    }
    switch (t) {
      case kCtTrue: case kCtFalse: if (elem) byte(); return;
      case kCtByte: byte(); return;
      case kCtI16: case kCtI32: case kCtI64: varint(); return;
      case kCtDouble: skip_bytes(8); return;
      case kCtBinary: skip_bytes(varint()); return;
      case kCtList: case kCtSet: {
        const uint8_t h = byte();
        uint64_t size = h >> 4;
        if (size == 15) size = varint();
        for (uint64_t i = 0; i < size && !bad; ++i) skip(h & 0x0F, depth + 1, true);
        return;
      }
      case kCtMap: {
        const uint64_t size = varint();
        if (!size) return;
        const uint8_t kv = byte();
        for (uint64_t i = 0; i < size && !bad; ++i) {
          skip(kv >> 4, depth + 1, true);
          skip(kv & 0x0F, depth + 1, true);
        }
        return;
      }
      case kCtStruct: {
        int16_t id = 0;
        int ft;
        while (field(&id, &ft) && !bad) skip(ft, depth + 1);
        return;
      }
      default: bad = true;
    }
  }
  // next field header of a struct; false at its stop byte
  bool field(int16_t* id, int* type) {
    const uint8_t h = byte();
    if (bad || h == kCtStop) return false;
    *type = h & 0x0F;
    const int d = h >> 4;
    *id = d ? (int16_t)(*id + d) : (int16_t)zigzag();
    return true;
  }
  int64_t integer(int t) {
    if (t == kCtI16 || t == kCtI32 || t == kCtI64) return zigzag();
    if (t == kCtByte) return (int8_t)byte();
    bad = true;
    return 0;
  }
};

enum : int { kPageData = 0, kPageIndex = 1, kPageDict = 2, kPageDataV2 = 3 };
enum : int { kEncPlain = 0, kEncPlainDict = 2, kEncRle = 3, kEncRleDict = 8 };

// The fields of a PageHeader the staging uses (parquet.thrift: PageHeader,
// DataPageHeader, DictionaryPageHeader, DataPageHeaderV2).
struct PageHeader {
  int32_t type = -1;
  int64_t usize = -1, csize = -1;
  bool has_v1 = false, has_dict = false, has_v2 = false;
  int64_t num_values = 0;
  int32_t encoding = -1, def_encoding = kEncRle;
  int64_t dict_n = 0;
  int32_t dict_encoding = kEncPlain;
  int64_t dl = 0, rl = 0;
  bool v2_compressed = true;
};

// Parse one header from [p, end): bytes it takes, 0 when [p, end) is too
// short for it, -1 when it is malformed.
int64_t parse_page_header(const uint8_t* p, const uint8_t* end, PageHeader* h) {
  Compact r{p, end};
  int16_t id = 0;
  int t;
  while (r.field(&id, &t)) {
    if (id == 1) h->type = (int32_t)r.integer(t);
    else if (id == 2) h->usize = r.integer(t);
    else if (id == 3) h->csize = r.integer(t);
    else if ((id == 5 || id == 7 || id == 8) && t == kCtStruct) {
      int16_t sid = 0;
      int st;
      if (id == 5) h->has_v1 = true;
      if (id == 7) h->has_dict = true;
      if (id == 8) h->has_v2 = true;
      while (r.field(&sid, &st) && !r.bad) {
        if (id == 5) {  // DataPageHeader
          if (sid == 1) h->num_values = r.integer(st);
          else if (sid == 2) h->encoding = (int32_t)r.integer(st);
          else if (sid == 3) h->def_encoding = (int32_t)r.integer(st);
          else r.skip(st, 1);
        } else if (id == 7) {  // DictionaryPageHeader
          if (sid == 1) h->dict_n = r.integer(st);
          else if (sid == 2) h->dict_encoding = (int32_t)r.integer(st);
          else r.skip(st, 1);
        } else {  // DataPageHeaderV2
          if (sid == 1) h->num_values = r.integer(st);
          else if (sid == 4) h->encoding = (int32_t)r.integer(st);
          else if (sid == 5) h->dl = r.integer(st);
          else if (sid == 6) h->rl = r.integer(st);
          else if (sid == 7 && (st == kCtTrue || st == kCtFalse)) h->v2_compressed = st == kCtTrue;
          else r.skip(st, 1);
        }
      }
    } else {
      r.skip(t, 0);
    }
    if (r.bad) break;
  }
  if (r.bad) return r.p >= end ? 0 : -1;
  return r.p - p;
}

// ----------------------------------------------------------------- staging

// One page as walked: where its body is in the chunk, and where its streams
// go in the chunk's staged region.
struct PagePlan {
  PageHeader h;
  uint64_t body = 0;     // chunk offset of the body
  uint64_t region = 0;   // region offset of the page's first stream
  uint64_t region2 = 0;  // v2: region offset of the values (the levels go to `region`)
};

struct ChunkPlan {
  std::vector<PagePlan> pages;
  uint64_t region_bytes = 0;  // staged bytes, pages laid end to end
  uint64_t offset = 0;        // of the region in the staged buffer
  // results of staging
  std::vector<wg_pq_page> out;  // stream pointers = region offsets, run pointers = offsets into runs
  std::vector<uint32_t> runs;
  uint64_t run_offset = 0;  // of `runs` in the joint run table, bytes
  uint64_t stream_bytes = 0;
};

bool read_at(int fd, uint8_t* dst, uint64_t n, uint64_t off) {
  while (n) {
    const ssize_t r = pread(fd, dst, n, (off_t)off);
    if (r < 0 && errno == EINTR) continue;
    if (r <= 0) return false;
    dst += r;
    n -= (uint64_t)r;
    off += (uint64_t)r;
  }
  return true;
}

struct Column {
  int fd = -1;
  const wg_pq_chunk* chunks = nullptr;
  int32_t phys = 0, max_def = 0;
  uint64_t width = 8;
};

// Walk the page headers of chunk c and plan its region.
wg_status walk_chunk(const Column& col, uint64_t c, ChunkPlan* plan, Failure& F) {
  const wg_pq_chunk& ch = col.chunks[c];
  std::vector<uint8_t> win;
  uint64_t pos = 0, rows = 0, region = 0;
  bool have_dict = false;
  while (pos < ch.bytes && rows < ch.num_rows) {
    if (F.failed.load()) return WG_ERR_INVALID;
    PageHeader h;
    int64_t used = 0;
    for (uint64_t w = 256;; w *= 4) {
      const uint64_t take = std::min<uint64_t>(w, ch.bytes - pos);
      win.resize(take);
      if (!read_at(col.fd, win.data(), take, ch.offset + pos))
        PQ_FAIL(F, WG_ERR_INVALID, "chunk %llu: read of %llu bytes at file offset %llu failed", (unsigned long long)c,
                (unsigned long long)take, (unsigned long long)(ch.offset + pos));
      h = PageHeader();
      used = parse_page_header(win.data(), win.data() + take, &h);
      if (used != 0 || take == ch.bytes - pos) break;
    }
    if (used <= 0 || h.type < 0 || h.usize < 0 || h.csize < 0)
      PQ_FAIL(F, WG_ERR_INVALID, "chunk %llu: malformed page header at chunk offset %llu", (unsigned long long)c,
              (unsigned long long)pos);
    PagePlan P;
    P.h = h;
    P.body = pos + (uint64_t)used;
    if (P.body + (uint64_t)h.csize > ch.bytes)
      PQ_FAIL(F, WG_ERR_INVALID, "chunk %llu: page body past the chunk's %llu bytes", (unsigned long long)c,
              (unsigned long long)ch.bytes);
    pos = P.body + (uint64_t)h.csize;
    const uint64_t usize = (uint64_t)h.usize;
    if (h.type == kPageDict) {
      if (!h.has_dict) PQ_FAIL(F, WG_ERR_INVALID, "chunk %llu: dictionary page without its header", (unsigned long long)c);
      if (h.dict_encoding != kEncPlain && h.dict_encoding != kEncPlainDict)
        PQ_FAIL(F, WG_ERR_UNSUPPORTED, "dictionary page encoding %d not on the GPU decode path", h.dict_encoding);
      if (h.dict_n < 0 || (uint64_t)h.dict_n * col.width > usize)
        PQ_FAIL(F, WG_ERR_INVALID, "dictionary page shorter than its values");
      P.region = region;
      region += align8(usize + 16);
      have_dict = true;
    } else if (h.type == kPageData || h.type == kPageDataV2) {
      const bool v2 = h.type == kPageDataV2;
      if (v2 ? !h.has_v2 : !h.has_v1)
        PQ_FAIL(F, WG_ERR_INVALID, "chunk %llu: data page without its header", (unsigned long long)c);
      if (h.num_values < 0 || h.num_values >= (1ll << 31))
        PQ_FAIL(F, WG_ERR_UNSUPPORTED, "data page of %lld values", (long long)h.num_values);
      if (h.encoding != kEncPlain && h.encoding != kEncPlainDict && h.encoding != kEncRleDict)
        PQ_FAIL(F, WG_ERR_UNSUPPORTED, "parquet encoding %d not on the GPU decode path", h.encoding);
      if (h.encoding != kEncPlain && !have_dict)
        PQ_FAIL(F, WG_ERR_INVALID, "dictionary-encoded page before a dictionary page");
      if (!v2 && col.max_def > 0 && h.def_encoding != kEncRle)
        PQ_FAIL(F, WG_ERR_UNSUPPORTED, "definition level encoding %d not on the GPU decode path", h.def_encoding);
      if (v2) {
        if (h.dl < 0 || h.rl < 0 || (uint64_t)(h.dl + h.rl) > usize || (uint64_t)(h.dl + h.rl) > (uint64_t)h.csize)
          PQ_FAIL(F, WG_ERR_INVALID, "v2 page levels longer than the page");
        P.region = region;
        if (col.max_def > 0 && h.dl > 0) region += align8((uint64_t)h.dl + 16);
        P.region2 = region;
        region += align8(usize - (uint64_t)(h.dl + h.rl) + 16);
      } else {
        P.region = region;
        region += align8(usize + 8 + 16);  // + 8: room to move PLAIN values up to an 8-byte boundary
      }
      rows += (uint64_t)h.num_values;
    } else {
      continue;  // index pages carry no values
    }
    plan->pages.push_back(P);
  }
  if (rows != ch.num_rows)
    PQ_FAIL(F, WG_ERR_INVALID, "column chunk pages hold %llu rows, the row group %llu", (unsigned long long)rows,
            (unsigned long long)ch.num_rows);
  plan->region_bytes = region;
  return WG_OK;
}

// Stage chunk c: its pages decompressed into dst (the chunk's region),
// streams scanned; the page table entries hold region / run-table offsets.
wg_status stage_chunk(const Column& col, uint64_t c, ChunkPlan* plan, uint8_t* dst, std::vector<uint8_t>* raw,
                      Failure& F) {
  const wg_pq_chunk& ch = col.chunks[c];
  raw->resize(ch.bytes + 8);
  if (!read_at(col.fd, raw->data(), ch.bytes, ch.offset))
    PQ_FAIL(F, WG_ERR_INVALID, "chunk %llu: read of %llu bytes at file offset %llu failed", (unsigned long long)c,
            (unsigned long long)ch.bytes, (unsigned long long)ch.offset);
  const uint8_t* R = raw->data();
  const uint64_t width = col.width;
  std::string why;
  int64_t dict_region = -1;
  uint64_t dict_n = 0;
  plan->out.clear();
  plan->runs.clear();
  plan->stream_bytes = 0;
  auto run_bytes = [&] { return (uint64_t)plan->runs.size() * 4; };
  for (const PagePlan& P : plan->pages) {
    if (F.failed.load()) return WG_ERR_INVALID;
    const PageHeader& h = P.h;
    const uint64_t usize = (uint64_t)h.usize, csize = (uint64_t)h.csize;
    if (h.type == kPageDict) {
      why.clear();
      const wg_status s = decompress(ch.codec, R + P.body, csize, dst + P.region, usize, &why);
      if (s != WG_OK) PQ_FAIL(F, s, "chunk %llu dictionary page: %s", (unsigned long long)c, why.c_str());
      dict_region = (int64_t)P.region;
      dict_n = (uint64_t)h.dict_n;
      plan->stream_bytes += dict_n * width;
      continue;
    }
    const uint64_t nv = (uint64_t)h.num_values;
    wg_pq_page pg;
    memset(&pg, 0, sizeof(pg));
    pg.num_rows = nv;
    pg.physical_type = col.phys;
    pg.encoding = h.encoding;
    const uint8_t* levels = nullptr;
    uint64_t levels_len = 0, levels_region = 0;  // region offset of the levels' first byte
    uint64_t values_region = 0, values_len = 0;
    if (h.type == kPageData) {
      why.clear();
      const wg_status s = decompress(ch.codec, R + P.body, csize, dst + P.region, usize, &why);
      if (s != WG_OK) PQ_FAIL(F, s, "chunk %llu data page: %s", (unsigned long long)c, why.c_str());
      uint64_t off = 0;
      if (col.max_def > 0) {
        if (usize < 4) PQ_FAIL(F, WG_ERR_INVALID, "data page too short for its levels");
        uint32_t ll;
        memcpy(&ll, dst + P.region, 4);
        if (4 + (uint64_t)ll > usize) PQ_FAIL(F, WG_ERR_INVALID, "definition levels longer than the page");
        levels = dst + P.region + 4;
        levels_len = ll;
        levels_region = P.region + 4;
        off = 4 + (uint64_t)ll;
      }
      values_region = P.region + off;
      values_len = usize - off;
    } else {  // v2: levels stored plain before the (maybe compressed) values
      const uint64_t dl = (uint64_t)h.dl, rl = (uint64_t)h.rl;
      if (col.max_def > 0 && dl > 0) {
        memcpy(dst + P.region, R + P.body + rl, dl);
        levels = dst + P.region;
        levels_len = dl;
        levels_region = P.region;
      }
      values_region = P.region2;
      values_len = usize - dl - rl;
      why.clear();
      const int codec = h.v2_compressed ? ch.codec : kCodecNone;
      const wg_status s = decompress(codec, R + P.body + rl + dl, csize - rl - dl, dst + values_region, values_len, &why);
      if (s != WG_OK) PQ_FAIL(F, s, "chunk %llu data page v2: %s", (unsigned long long)c, why.c_str());
    }
    uint64_t present = nv;
    if (levels) {
      // read in place from the 8-byte word at or below the levels
      const uint32_t bias = (uint32_t)(levels_region & 7) * 8;
      const uint64_t r0 = run_bytes();
      uint64_t used = 0, ones = 0;
      why.clear();
      const wg_status s = scan_runs(levels, levels_len, 1, nv, bias, &plan->runs, &used, &ones, &why);
      if (s != WG_OK) PQ_FAIL(F, s, "chunk %llu definition levels: %s", (unsigned long long)c, why.c_str());
      present = ones;
      if (present < nv) {  // a page without NULLs decodes as a required one
        pg.def.runs = (const uint32_t*)(uintptr_t)r0;
        pg.def.n_runs = (uint32_t)((run_bytes() - r0) / 16);
        pg.def.bit_width = 1;
        pg.def.data = (const uint8_t*)(uintptr_t)(levels_region & ~7ull);
        plan->stream_bytes += levels_len + (run_bytes() - r0);
      } else {
        plan->runs.resize(r0 / 4);
      }
    }
    if (h.encoding == kEncPlain) {
      const uint64_t need = present * width;
      if (values_len < need) PQ_FAIL(F, WG_ERR_INVALID, "PLAIN page shorter than its values");
      uint64_t at = values_region;
      if (at & 7) {  // follows levels: move up to the next 8-byte boundary (room planned)
        memmove(dst + align8(at), dst + at, need);
        at = align8(at);
      }
      pg.plain = (const void*)(uintptr_t)at;
      plan->stream_bytes += need;
    } else {
      const uint64_t r0 = run_bytes();
      int bw = 0;
      if (present > 0) {
        if (values_len < 1) PQ_FAIL(F, WG_ERR_INVALID, "dictionary index stream without its bit width");
        bw = dst[values_region];
        const uint64_t s_region = values_region + 1;
        uint64_t used = 0;
        why.clear();
        const wg_status s = scan_runs(dst + s_region, values_len - 1, bw, present, (uint32_t)(s_region & 7) * 8,
                                      &plan->runs, &used, nullptr, &why);
        if (s != WG_OK) PQ_FAIL(F, s, "chunk %llu dictionary indices: %s", (unsigned long long)c, why.c_str());
        plan->stream_bytes += (values_len - 1) + (run_bytes() - r0);
      }
      pg.idx.runs = (const uint32_t*)(uintptr_t)r0;
      pg.idx.n_runs = (uint32_t)((run_bytes() - r0) / 16);
      pg.idx.bit_width = bw;
      pg.idx.data = (const uint8_t*)(uintptr_t)((values_region + 1) & ~7ull);
      pg.dict = (const void*)(uintptr_t)dict_region;
      pg.dict_n = dict_n;
    }
    plan->out.push_back(pg);
  }
  return WG_OK;
}

int usable_threads(int32_t want) {
  int n = 0;
  cpu_set_t set;
  if (sched_getaffinity(0, sizeof(set), &set) == 0) n = CPU_COUNT(&set);
  if (n <= 0) n = (int)std::thread::hardware_concurrency();
  n = std::max(1, std::min(n, 16));
  return want > 0 ? std::min<int>(want, 64) : n;
}

// fn(worker, item) for items [0, n) on `workers` threads (claimed in order).
template <class Fn>
void parallel_for(uint64_t n, int workers, Fn fn) {
  std::atomic<uint64_t> next{0};
  auto body = [&](int w) {
    for (uint64_t i; (i = next.fetch_add(1)) < n;) fn(w, i);
  };
  if (workers <= 1 || n <= 1) {
    body(0);
    return;
  }
  std::vector<std::thread> pool;
  for (int w = 1; w < workers; ++w) pool.emplace_back(body, w);
  body(0);
  for (auto& t : pool) t.join();
}

constexpr uint64_t kSlotBudget = 1ull << 30;  // pinned staging slots, all threads together

}  // namespace
}  // namespace wg

struct wg_pq_staged {
  wg_ctx* ctx = nullptr;  // null: host memory
  uint8_t* streams = nullptr;
  uint8_t* runs = nullptr;
  uint64_t stream_region = 0, run_region = 0;
  uint64_t rows = 0, stream_bytes = 0;
  std::vector<wg_pq_page> pages;
};

using namespace wg;

namespace {

void release(wg_pq_staged* s) {
  if (!s) return;
  if (s->ctx) {
    if (s->streams) (void)hipFree(s->streams);
    if (s->runs) (void)hipFree(s->runs);
  } else {
    free(s->streams);
    free(s->runs);
  }
  delete s;
}

struct Fd {
  int fd = -1;
  ~Fd() {
    if (fd >= 0) close(fd);
  }
};

}  // namespace

extern "C" {

wg_status wg_pq_decompress(int32_t codec, const uint8_t* src, uint64_t len, uint8_t* dst, uint64_t dst_len) {
  clear_error();
  WG_CHECK((src || len == 0) && (dst || dst_len == 0), WG_ERR_INVALID, "null buffer");
  std::string why;
  const wg_status s = decompress(codec, src, len, dst, dst_len, &why);
  if (s != WG_OK) set_error("%s", why.c_str());
  return s;
}

wg_status wg_pq_scan_runs(const uint8_t* data, uint64_t len, int32_t bit_width, uint64_t count, uint32_t* runs,
                          uint64_t cap, uint64_t* n_runs, uint64_t* used, uint64_t* ones) {
  clear_error();
  WG_CHECK(data || len == 0, WG_ERR_INVALID, "null data");
  WG_CHECK(n_runs && used, WG_ERR_INVALID, "null argument");
  std::vector<uint32_t> table;
  std::string why;
  uint64_t n1 = 0;
  const wg_status s = scan_runs(data, len, bit_width, count, 0, &table, used, &n1, &why);
  if (s != WG_OK) {
    set_error("%s", why.c_str());
    return s;
  }
  const uint64_t nr = table.size() / 4;
  *n_runs = nr;
  if (ones) *ones = n1;
  if (runs) {
    if (nr && cap) memcpy(runs, table.data(), std::min(nr, cap) * 16);
    WG_CHECK(nr <= cap, WG_ERR_INVALID, "%llu runs do not fit %llu", (unsigned long long)nr, (unsigned long long)cap);
  }
  return WG_OK;
}

wg_status wg_pq_stage(wg_ctx* ctx, const char* path, const wg_pq_chunk* chunks, uint64_t n_chunks,
                      int32_t physical_type, int32_t max_def, int32_t threads, wg_pq_staged** out) {
  clear_error();
  WG_CHECK(path && out && (chunks || n_chunks == 0), WG_ERR_INVALID, "null argument");
  *out = nullptr;
  WG_CHECK(physical_type == 1 || physical_type == 2 || physical_type == 5, WG_ERR_UNSUPPORTED,
           "parquet physical type %d not on the GPU decode path", physical_type);
  WG_CHECK(max_def == 0 || max_def == 1, WG_ERR_UNSUPPORTED, "definition level %d not on the GPU decode path",
           max_def);
  for (uint64_t c = 0; c < n_chunks; ++c) {
    const int codec = chunks[c].codec;
    WG_CHECK(codec == kCodecNone || codec == kCodecSnappy || codec == kCodecGzip || codec == kCodecLz4 ||
                 codec == kCodecZstd || codec == kCodecLz4Raw,
             WG_ERR_UNSUPPORTED, "compression codec %d not on the GPU decode path", codec);
    WG_CHECK(chunks[c].bytes < (1ull << 40), WG_ERR_INVALID, "chunk of %llu bytes",
             (unsigned long long)chunks[c].bytes);
  }
  Fd fd;
  fd.fd = open(path, O_RDONLY | O_CLOEXEC);
  WG_CHECK(fd.fd >= 0, WG_ERR_INVALID, "cannot open %s: %s", path, strerror(errno));
  Column col;
  col.fd = fd.fd;
  col.chunks = chunks;
  col.phys = physical_type;
  col.max_def = max_def;
  col.width = physical_type == 1 ? 4 : 8;
  const int nthreads = usable_threads(threads);

  // walk: every chunk's pages and region
  std::vector<ChunkPlan> plan(n_chunks);
  Failure F;
  parallel_for(n_chunks, std::min<int>(nthreads, (int)std::max<uint64_t>(n_chunks, 1)),
               [&](int, uint64_t c) { (void)walk_chunk(col, c, &plan[c], F); });
  WG_TRY(F.report());
  uint64_t total = 0, slot = 0;
  for (auto& p : plan) {
    p.offset = total;
    total += p.region_bytes;
    slot = std::max(slot, p.region_bytes);
  }
  slot = (slot + 255) & ~255ull;

  std::unique_ptr<wg_pq_staged, void (*)(wg_pq_staged*)> S(new wg_pq_staged, release);
  S->ctx = ctx;
  int workers = (int)std::min<uint64_t>((uint64_t)nthreads, std::max<uint64_t>(n_chunks, 1));
  std::vector<std::vector<uint8_t>> raw(workers);
  if (!ctx) {
    S->streams = (uint8_t*)aligned_alloc(256, (total + 64 + 255) & ~255ull);
    WG_CHECK(S->streams, WG_ERR_OOM, "host staging of %llu bytes", (unsigned long long)total);
    parallel_for(n_chunks, workers, [&](int w, uint64_t c) {
      (void)stage_chunk(col, c, &plan[c], S->streams + plan[c].offset, &raw[w], F);
    });
    WG_TRY(F.report());
  } else {
    WG_HIP(hipSetDevice(ctx->device));
    WG_TRY(dev_alloc(ctx, (void**)&S->streams, total + 64, nullptr, "staged Parquet pages"));
    workers = std::max(1, (int)std::min<uint64_t>((uint64_t)workers, std::max<uint64_t>(kSlotBudget / std::max<uint64_t>(slot, 1), 1)));
    std::lock_guard<std::mutex> upload_lock(ctx->upload_mu);  // the slots live in the context's pinned buffer
    const uint64_t need = slot * (uint64_t)workers;
    if (ctx->upload_cap < need) {
      if (ctx->upload) WG_HIP(hipHostFree(ctx->upload));
      ctx->upload = nullptr;
      ctx->upload_cap = 0;
      WG_CHECK(hipHostMalloc(&ctx->upload, need, hipHostMallocDefault) == hipSuccess, WG_ERR_OOM,
               "hipHostMalloc(%llu) for the staging slots failed", (unsigned long long)need);
      ctx->upload_cap = need;
    }
    std::vector<hipEvent_t> done(workers, nullptr);
    std::vector<char> pending(workers, 0);
    for (auto& e : done) WG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    const int device = ctx->device;
    parallel_for(n_chunks, workers, [&](int w, uint64_t c) {
      if (F.failed.load()) return;
      if (hipSetDevice(device) != hipSuccess) {
        F.set(WG_ERR_DEVICE, "hipSetDevice(%d) on a staging thread failed", device);
        return;
      }
      uint8_t* sl = (uint8_t*)ctx->upload + (uint64_t)w * slot;
      if (pending[w] && hipEventSynchronize(done[w]) != hipSuccess) {
        F.set(WG_ERR_DEVICE, "staging copy failed");
        return;
      }
      pending[w] = 0;
      if (stage_chunk(col, c, &plan[c], sl, &raw[w], F) != WG_OK) return;
      if (hipMemcpyAsync(S->streams + plan[c].offset, sl, plan[c].region_bytes, hipMemcpyHostToDevice,
                         ctx->stream) != hipSuccess ||
          hipEventRecord(done[w], ctx->stream) != hipSuccess) {
        F.set(WG_ERR_DEVICE, "staging copy of chunk %llu failed", (unsigned long long)c);
        return;
      }
      pending[w] = 1;
    });
    const hipError_t sync = hipStreamSynchronize(ctx->stream);
    for (auto& e : done) (void)hipEventDestroy(e);
    WG_TRY(F.report());
    WG_CHECK(sync == hipSuccess, WG_ERR_DEVICE, "staging copies: %s", hipGetErrorString(sync));
  }

  // the run tables end to end, then the page table with final addresses
  uint64_t run_total = 0;
  for (auto& p : plan) {
    p.run_offset = run_total;
    run_total += (uint64_t)p.runs.size() * 4;
  }
  const uint64_t run_alloc = run_total + 64;
  std::vector<uint8_t> joined(run_alloc, 0);
  for (auto& p : plan)
    if (!p.runs.empty()) memcpy(joined.data() + p.run_offset, p.runs.data(), p.runs.size() * 4);
  if (ctx) {
    WG_TRY(dev_alloc(ctx, (void**)&S->runs, run_alloc, nullptr, "staged Parquet run tables"));
    WG_HIP(hipMemcpy(S->runs, joined.data(), run_alloc, hipMemcpyHostToDevice));
  } else {
    S->runs = (uint8_t*)aligned_alloc(256, (run_alloc + 255) & ~255ull);
    WG_CHECK(S->runs, WG_ERR_OOM, "host run tables of %llu bytes", (unsigned long long)run_alloc);
    memcpy(S->runs, joined.data(), run_alloc);
  }
  S->stream_region = total;
  S->run_region = run_total;
  for (auto& p : plan) {
    uint8_t* sb = S->streams + p.offset;
    uint8_t* rb = S->runs + p.run_offset;
    for (wg_pq_page pg : p.out) {
      if (pg.def.n_runs) {
        pg.def.runs = (const uint32_t*)(rb + (uintptr_t)pg.def.runs);
        pg.def.data = sb + (uintptr_t)pg.def.data;
      }
      if (pg.encoding == kEncPlain) {
        pg.plain = sb + (uintptr_t)pg.plain;
      } else {
        pg.idx.runs = (const uint32_t*)(rb + (uintptr_t)pg.idx.runs);
        pg.idx.data = sb + (uintptr_t)pg.idx.data;
        pg.dict = sb + (uintptr_t)pg.dict;
      }
      S->rows += pg.num_rows;
      S->pages.push_back(pg);
    }
    S->stream_bytes += p.stream_bytes;
  }
  *out = S.release();
  return WG_OK;
}

wg_status wg_pq_staged_info(const wg_pq_staged* s, uint64_t* rows, uint64_t* n_pages, uint64_t* stream_bytes,
                            uint64_t* staged_bytes) {
  clear_error();
  WG_CHECK(s, WG_ERR_INVALID, "null staged column");
  if (rows) *rows = s->rows;
  if (n_pages) *n_pages = s->pages.size();
  if (stream_bytes) *stream_bytes = s->stream_bytes;
  if (staged_bytes) *staged_bytes = s->stream_region + s->run_region;
  return WG_OK;
}

wg_status wg_pq_staged_pages(const wg_pq_staged* s, wg_pq_page* pages, uint64_t cap) {
  clear_error();
  WG_CHECK(s && (pages || cap == 0), WG_ERR_INVALID, "null argument");
  const uint64_t n = std::min<uint64_t>(cap, s->pages.size());
  if (n) memcpy(pages, s->pages.data(), n * sizeof(wg_pq_page));
  return WG_OK;
}

wg_status wg_pq_staged_free(wg_pq_staged* s) {
  clear_error();
  if (s && s->ctx) (void)hipSetDevice(s->ctx->device);
  release(s);
  return WG_OK;
}

}  // extern "C"

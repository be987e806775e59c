// Robustness of the Parquet host staging against malformed input, in the
// manner of the reference's fuzz tests (internal/storage/parquet/
// page_reader_fuzz_test.go: FuzzRLEDecode, FuzzDecodeBitPacked;
// thrift_fuzz_test.go: FuzzDecodePageHeader): built with AddressSanitizer and
// UBSan, this program stages mutated copies of a valid file (bytes flipped,
// set, zeroed or cut short inside its column chunks) and feeds random bytes to
// the hybrid run scanner and the decompressors. Every call must return, with
// WG_OK or an error status; any out-of-bounds access aborts the run.
//
// usage: pq_fuzz <file> <physical type> <max def> <iterations> <seed> <tmp file>
//                <n chunks> {<offset> <bytes> <rows> <codec>}...
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "../../include/wadjet_gpu.h"

static std::vector<uint8_t> read_file(const char* path) {
  std::vector<uint8_t> out;
  FILE* f = std::fopen(path, "rb");
  if (!f) return out;
  uint8_t buf[1 << 16];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) out.insert(out.end(), buf, buf + n);
  std::fclose(f);
  return out;
}

static bool write_file(const char* path, const std::vector<uint8_t>& b) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return false;
  const bool ok = std::fwrite(b.data(), 1, b.size(), f) == b.size();
  return std::fclose(f) == 0 && ok;
}

int main(int argc, char** argv) {
  if (argc < 8) {
    std::fprintf(stderr, "usage: pq_fuzz file phys maxdef iterations seed tmp nchunks {off bytes rows codec}...\n");
    return 2;
  }
  const char* path = argv[1];
  const int32_t phys = std::atoi(argv[2]), max_def = std::atoi(argv[3]);
  const int iters = std::atoi(argv[4]);
  std::mt19937_64 rng(std::strtoull(argv[5], nullptr, 10));
  const char* tmp = argv[6];
  const int nc = std::atoi(argv[7]);
  if (argc != 8 + 4 * nc) return 2;
  std::vector<wg_pq_chunk> chunks(nc);
  for (int c = 0; c < nc; ++c) {
    chunks[c].offset = std::strtoull(argv[8 + 4 * c], nullptr, 10);
    chunks[c].bytes = std::strtoull(argv[9 + 4 * c], nullptr, 10);
    chunks[c].num_rows = std::strtoull(argv[10 + 4 * c], nullptr, 10);
    chunks[c].codec = std::atoi(argv[11 + 4 * c]);
    chunks[c].reserved = 0;
  }
  const std::vector<uint8_t> orig = read_file(path);
  if (orig.empty()) return 2;
  int ok = 0, err = 0;
  for (int it = 0; it < iters; ++it) {
    std::vector<uint8_t> b = orig;
    std::vector<wg_pq_chunk> ch = chunks;
    const int c = (int)(rng() % nc);
    const uint64_t lo = ch[c].offset, span = std::max<uint64_t>(ch[c].bytes, 1);
    const int kind = (int)(rng() % 5);
    const int edits = 1 + (int)(rng() % 8);
    for (int e = 0; e < edits; ++e) {
      const uint64_t at = lo + rng() % span;
      if (at >= b.size()) continue;
      switch (kind) {
        case 0: b[at] ^= (uint8_t)(1u << (rng() % 8)); break;                 // one bit
        case 1: b[at] = (uint8_t)rng(); break;                               // one byte
        case 2: b[at] = 0xFF; break;                                         // varint continuation
        case 3: b[at] = 0; break;
        default: ch[c].bytes = rng() % span; break;                          // the chunk cut short
      }
    }
    if (!write_file(tmp, b)) return 2;
    wg_pq_staged* s = nullptr;
    const wg_status st = wg_pq_stage(nullptr, tmp, ch.data(), (uint64_t)nc, phys, max_def, 1 + (int)(rng() % 3), &s);
    if (st == WG_OK) {
      ++ok;
      uint64_t rows = 0, pages = 0;
      wg_pq_staged_info(s, &rows, &pages, nullptr, nullptr);
      std::vector<wg_pq_page> table(pages ? pages : 1);
      wg_pq_staged_pages(s, table.data(), pages);
      wg_pq_staged_free(s);
    } else {
      ++err;
    }
  }
  // the run scanner and the decompressors on random bytes (FuzzRLEDecode, FuzzDecodeBitPacked)
  for (int it = 0; it < iters; ++it) {
    std::vector<uint8_t> data(rng() % 64);
    for (auto& x : data) x = (uint8_t)rng();
    const int bw = (int)(rng() % 34) - 1;
    const uint64_t count = rng() % 2000;
    uint64_t nr = 0, used = 0, ones = 0;
    if (wg_pq_scan_runs(data.empty() ? nullptr : data.data(), data.size(), bw, count, nullptr, 0, &nr, &used, &ones) ==
        WG_OK) {
      std::vector<uint32_t> runs(4 * (nr ? nr : 1));
      wg_pq_scan_runs(data.empty() ? nullptr : data.data(), data.size(), bw, count, runs.data(), nr, &nr, &used, &ones);
    }
    static const int codecs[] = {0, 1, 2, 5, 6, 7};
    const uint64_t out_len = rng() % 4096;
    std::vector<uint8_t> out(out_len ? out_len : 1);
    wg_pq_decompress(codecs[rng() % 6], data.empty() ? nullptr : data.data(), data.size(), out.data(), out_len);
  }
  std::printf("staged %d, refused %d\n", ok, err);
  return 0;
}

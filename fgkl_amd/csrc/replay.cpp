// GATK PairHMM / PD-HMM results dumps: reading them into batches with their recorded
// likelihoods, and writing batches back out in the same formats.
//
// Formats (what HaplotypeCaller writes with --pair-hmm-results-file / --pdhmm-results-file, as
// fgkl's replay tools read them: crates/pairhmm/src/bin/replay.rs:83-159,
// pdreplay.rs:34-115): one line per (read, haplotype), reads outer and haplotypes inner, region
// after region; '#' lines are comments.
//  * plain: space-separated `hap read quals ins_gop del_gop gcp log10L`, qualities Phred+33;
//  * PD: tab-separated `hap flags read quals ins_gop del_gop gcp log10L`, flags rendered by
//    Java's Arrays.toString(byte[]) ("[0, 4, -128]", signed bytes).
// The dump holds no region boundaries.  A region's haplotype list is the run of haplotypes its
// first read is paired with, known once the first haplotype comes round again; every later read
// repeats the list.  A read starts when the list starts over (duplicate reads sit back to back,
// so the read text never marks a boundary); a line that does not continue the list either starts
// a new region (after a complete read) or, mid-read, throws the partial region away.  Regions
// whose line count is not reads x haplotypes are dropped at the end.
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "fgkl_replay.hpp"

namespace fgkl {
namespace {

struct Hap {
  std::vector<uint8_t> bases, flags;
  bool operator==(const Hap& o) const { return bases == o.bases && flags == o.flags; }
};

struct PartialRegion {
  std::vector<Read> reads;
  std::vector<Hap> haps;
  std::vector<double> expected;
  void clear() {
    reads.clear();
    haps.clear();
    expected.clear();
  }
};

std::vector<uint8_t> phred33(const std::string& s) {
  std::vector<uint8_t> q(s.size());
  for (size_t i = 0; i < s.size(); ++i) q[i] = uint8_t(uint8_t(s[i]) - 33);
  return q;
}

std::vector<uint8_t> bytes_of(const std::string& s) { return std::vector<uint8_t>(s.begin(), s.end()); }

// "[0, 4, -128]" -> {0, 4, 128}
bool java_byte_array(const std::string& s, std::vector<uint8_t>& out) {
  out.clear();
  size_t i = 0;
  while (i < s.size() && (s[i] == '[' || s[i] == ' ')) ++i;
  while (i < s.size() && s[i] != ']') {
    char* end = nullptr;
    const long v = std::strtol(s.c_str() + i, &end, 10);
    if (end == s.c_str() + i || v < -128 || v > 255) return false;
    out.push_back(uint8_t(v & 0xFF));
    i = size_t(end - s.c_str());
    while (i < s.size() && (s[i] == ',' || s[i] == ' ')) ++i;
  }
  return true;
}

std::vector<std::string> split(const std::string& line, char sep) {
  std::vector<std::string> f;
  size_t a = 0;
  for (;;) {
    const size_t b = line.find(sep, a);
    f.push_back(line.substr(a, b == std::string::npos ? std::string::npos : b - a));
    if (b == std::string::npos) break;
    a = b + 1;
  }
  return f;
}

// Appends a region whose lines form reads x haplotypes; returns whether it did.
bool emit(const PartialRegion& r, ReplayData& out) {
  if (r.reads.empty() || r.expected.size() != r.reads.size() * r.haps.size()) {
    if (!r.reads.empty()) ++out.dropped_regions;
    return false;
  }
  out.batch.begin_region();
  for (const Read& rd : r.reads) out.batch.add_read(rd);
  for (const Hap& h : r.haps) out.batch.add_hap(h.bases, &h.flags);
  out.expected.insert(out.expected.end(), r.expected.begin(), r.expected.end());
  return true;
}

}  // namespace

int replay_read(const char* path, bool pd, int64_t max_regions, ReplayData& out,
                std::string& error) {
  std::ifstream in(path);
  if (!in) {
    error = std::string("cannot open ") + path + ": " + std::strerror(errno);
    return 1;
  }
  out = ReplayData{};
  const size_t nf = pd ? 8 : 7;
  const char sep = pd ? '\t' : ' ';
  PartialRegion cur;
  bool list_known = false;
  size_t next_hap = 0;
  int64_t complete = 0;
  std::string line;
  while (std::getline(in, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line.empty() || line[0] == '#') continue;
    ++out.lines;
    const std::vector<std::string> f = split(line, sep);
    Hap hap;
    if (f.size() < nf || (pd && !java_byte_array(f[1], hap.flags))) {
      ++out.skipped_lines;
      continue;
    }
    hap.bases = bytes_of(f[0]);
    if (!pd) hap.flags.assign(hap.bases.size(), 0);
    const size_t r0 = pd ? 2 : 1;
    char* end = nullptr;
    const double lk = std::strtod(f[nf - 1].c_str(), &end);
    if (end == f[nf - 1].c_str() || hap.flags.size() != hap.bases.size() ||
        f[r0].size() != f[r0 + 1].size() || f[r0].size() != f[r0 + 2].size() ||
        f[r0].size() != f[r0 + 3].size() || f[r0].size() != f[r0 + 4].size()) {
      ++out.skipped_lines;
      continue;
    }
    auto read = [&]() {
      Read r;
      r.bases = bytes_of(f[r0]);
      r.quals = phred33(f[r0 + 1]);
      r.ins = phred33(f[r0 + 2]);
      r.del = phred33(f[r0 + 3]);
      r.gcp = phred33(f[r0 + 4]);
      return r;
    };
    const bool starts_over = !cur.haps.empty() && cur.haps.front() == hap;
    if (cur.reads.empty()) {
      cur.reads.push_back(read());
      cur.haps.push_back(hap);
      next_hap = 1;
    } else if (!list_known) {
      if (starts_over) {  // the first read's list is complete
        list_known = true;
        cur.reads.push_back(read());
        next_hap = 1;
      } else {
        cur.haps.push_back(hap);
        ++next_hap;
      }
    } else if (next_hap == cur.haps.size()) {  // after a complete read
      if (starts_over) {
        cur.reads.push_back(read());
        next_hap = 1;
      } else {
        if (emit(cur, out)) ++complete;
        if (max_regions > 0 && complete >= max_regions) {
          cur.clear();
          break;
        }
        cur.clear();
        cur.reads.push_back(read());
        cur.haps.push_back(hap);
        list_known = false;
        next_hap = 1;
      }
    } else if (cur.haps[next_hap] == hap) {
      ++next_hap;
    } else {  // the list breaks mid-read: drop what was gathered, start from this line
      ++out.dropped_regions;
      cur.clear();
      cur.reads.push_back(read());
      cur.haps.push_back(hap);
      list_known = false;
      next_hap = 1;
    }
    cur.expected.push_back(lk);
  }
  emit(cur, out);
  out.batch.end_batch();
  return 0;
}

int replay_write(const Batch& b, const std::vector<double>& expected, const char* path, bool pd,
                 std::string& error) {
  FILE* f = std::fopen(path, "w");
  if (!f) {
    error = std::string("cannot write ") + path + ": " + std::strerror(errno);
    return 1;
  }
  std::fprintf(f, pd ? "#hap\tflags\tread\tquals\tins\tdel\tgcp\tlog10L\n"
                     : "# hap-bases read-bases read-qual read-ins-qual read-del-qual gcp "
                       "expected-result\n");
  const char sep = pd ? '\t' : ' ';
  size_t e = 0;
  auto text = [](const uint8_t* p, int64_t n, int add) {
    std::string s(size_t(n), ' ');
    for (int64_t i = 0; i < n; ++i) s[size_t(i)] = char(p[i] + add);
    return s;
  };
  const int64_t G = int64_t(b.region_reads.size()) - 1;
  for (int64_t g = 0; g < G; ++g) {
    for (int64_t r = b.region_reads[size_t(g)]; r < b.region_reads[size_t(g + 1)]; ++r) {
      const int64_t o = b.read_offsets[size_t(r)], n = b.read_offsets[size_t(r + 1)] - o;
      const std::string rd = text(b.bases.data() + o, n, 0) + sep +
                             text(b.quals.data() + o, n, 33) + sep +
                             text(b.ins.data() + o, n, 33) + sep +
                             text(b.del.data() + o, n, 33) + sep + text(b.gcp.data() + o, n, 33);
      for (int64_t h = b.region_haps[size_t(g)]; h < b.region_haps[size_t(g + 1)]; ++h) {
        const int64_t ho = b.hap_offsets[size_t(h)], m = b.hap_offsets[size_t(h + 1)] - ho;
        std::string line = text(b.hap_bases.data() + ho, m, 0);
        if (pd) {
          line += "\t[";
          for (int64_t i = 0; i < m; ++i) {
            if (i) line += ", ";
            line += std::to_string(int(int8_t(b.hap_flags[size_t(ho + i)])));
          }
          line += "]";
        }
        const double v = e < expected.size() ? expected[e++] : 0.0;
        std::fprintf(f, "%s%c%s%c%.17e\n", line.c_str(), sep, rd.c_str(), sep, v);
      }
    }
  }
  const bool ok = std::fclose(f) == 0;
  if (!ok) error = std::string("cannot write ") + path;
  return ok ? 0 : 1;
}

}  // namespace fgkl

// Deterministic synthetic regions (synthetic.rs:1-334 restated in C++, draw for draw, so the
// reference's own test regions regenerate bit-identically), plus the generators of the five
// benchmark workloads, which follow the same SplitMix64 stream conventions.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "fgkl_synthetic.hpp"

namespace fgkl {

namespace {
constexpr uint8_t kBases[4] = {'A', 'C', 'G', 'T'};
}

// synthetic.rs:19-25
uint64_t Rng::next_u64() {
  s_ += 0x9E3779B97F4A7C15ull;
  uint64_t z = s_;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
// synthetic.rs:28-30
uint64_t Rng::below(uint64_t n) { return next_u64() % n; }
// synthetic.rs:33-35
bool Rng::chance(double p) { return double(next_u64() >> 11) / double(uint64_t(1) << 53) < p; }
// synthetic.rs:38-40
uint8_t Rng::base() { return kBases[below(4)]; }
// Inclusive integer range.
int64_t Rng::range(int64_t lo, int64_t hi) { return lo + int64_t(below(uint64_t(hi - lo + 1))); }
// Standard normal by Box-Muller from two draws; u1 lies in (0, 1].
double Rng::normal() {
  const double u1 = double((next_u64() >> 11) + 1) / double(uint64_t(1) << 53);
  const double u2 = double(next_u64() >> 11) / double(uint64_t(1) << 53);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586476925286766559 * u2);
}

void Batch::begin_region() {
  region_reads.push_back(int64_t(read_offsets.size()) - 1);
  region_haps.push_back(int64_t(hap_offsets.size()) - 1);
}
void Batch::end_batch() {
  region_reads.push_back(int64_t(read_offsets.size()) - 1);
  region_haps.push_back(int64_t(hap_offsets.size()) - 1);
}
void Batch::add_read(const Read& r) {
  bases.insert(bases.end(), r.bases.begin(), r.bases.end());
  quals.insert(quals.end(), r.quals.begin(), r.quals.end());
  ins.insert(ins.end(), r.ins.begin(), r.ins.end());
  del.insert(del.end(), r.del.begin(), r.del.end());
  gcp.insert(gcp.end(), r.gcp.begin(), r.gcp.end());
  read_offsets.push_back(int64_t(bases.size()));
}
void Batch::add_hap(const std::vector<uint8_t>& b, const std::vector<uint8_t>* f) {
  hap_bases.insert(hap_bases.end(), b.begin(), b.end());
  if (f)
    hap_flags.insert(hap_flags.end(), f->begin(), f->end());
  else
    hap_flags.insert(hap_flags.end(), b.size(), 0);
  hap_offsets.push_back(int64_t(hap_bases.size()));
}
int64_t Batch::cells() const {  // synthetic.rs:166-170, summed over regions
  int64_t total = 0;
  for (size_t g = 0; g + 1 < region_reads.size(); ++g) {
    const int64_t rb = read_offsets[size_t(region_reads[g + 1])] - read_offsets[size_t(region_reads[g])];
    const int64_t hb = hap_offsets[size_t(region_haps[g + 1])] - hap_offsets[size_t(region_haps[g])];
    total += rb * hb;
  }
  return total;
}

// synthetic.rs:312-334
Read sample_read(Rng& rng, const std::vector<uint8_t>& source, int64_t read_len) {
  int64_t len = read_len;
  if (rng.chance(0.3)) len = 1 + int64_t(rng.below(uint64_t(read_len)));
  len = std::min<int64_t>(len, int64_t(source.size()));
  const int64_t start = int64_t(rng.below(uint64_t(int64_t(source.size()) - len + 1)));
  Read r;
  r.bases.assign(source.begin() + start, source.begin() + start + len);
  for (auto& b : r.bases) {
    if (rng.chance(0.002))
      b = 'N';
    else if (rng.chance(0.01))
      b = rng.base();
    r.quals.push_back(uint8_t(rng.chance(0.05) ? 2 + rng.below(15) : 20 + rng.below(21)));
    r.ins.push_back(uint8_t(rng.chance(0.1) ? 20 + rng.below(30) : 45));
    r.del.push_back(uint8_t(rng.chance(0.1) ? 20 + rng.below(30) : 45));
  }
  r.gcp.assign(size_t(len), 10);
  return r;
}

// synthetic.rs:84-98
std::vector<uint8_t> edited(Rng& rng, const std::vector<uint8_t>& alphabet,
                            const std::vector<uint8_t>& base) {
  std::vector<uint8_t> hap = base;
  const uint64_t edits = 1 + rng.below(3);
  for (uint64_t e = 0; e < edits; ++e) {
    const uint64_t pos = rng.below(hap.size());
    switch (rng.below(3)) {
      case 0: hap[pos] = alphabet[rng.below(alphabet.size())]; break;
      case 1: hap.insert(hap.begin() + int64_t(pos), alphabet[rng.below(alphabet.size())]); break;
      default:
        if (hap.size() > 1) hap.erase(hap.begin() + int64_t(pos));
        break;
    }
  }
  return hap;
}

// synthetic.rs:103-153
void region_generate(Batch& out, uint64_t seed, int64_t num_reads, int64_t read_len,
                     int64_t num_haps, int64_t hap_len) {
  Rng rng(seed);
  std::vector<uint8_t> reference(static_cast<size_t>(hap_len));
  for (auto& b : reference) b = rng.base();
  std::vector<std::vector<uint8_t>> haps{reference};
  while (int64_t(haps.size()) < num_haps) {
    if (haps.size() > 1 && rng.chance(0.1)) {
      std::vector<uint8_t> dup = haps[rng.below(haps.size())];
      haps.push_back(dup);
      continue;
    }
    std::vector<uint8_t> hap = reference;
    const uint64_t edits = 1 + rng.below(3);
    for (uint64_t e = 0; e < edits; ++e) {
      const uint64_t pos = rng.below(hap.size());
      switch (rng.below(3)) {
        case 0: {
          const uint8_t old = hap[pos];
          uint8_t nb = rng.base();
          while (nb == old) nb = rng.base();
          hap[pos] = nb;
          break;
        }
        case 1: {
          const uint64_t n = 1 + rng.below(5);
          std::vector<uint8_t> ins(n);
          for (auto& b : ins) b = rng.base();
          hap.insert(hap.begin() + int64_t(pos), ins.begin(), ins.end());
          break;
        }
        default: {
          const uint64_t end = std::min<uint64_t>(pos + 1 + rng.below(5), hap.size());
          if (hap.size() - (end - pos) >= 1) hap.erase(hap.begin() + int64_t(pos), hap.begin() + int64_t(end));
          break;
        }
      }
    }
    haps.push_back(hap);
  }
  out.begin_region();
  for (int64_t i = 0; i < num_reads; ++i) {
    const auto& hap = haps[rng.below(haps.size())];
    out.add_read(sample_read(rng, hap, read_len));
  }
  for (const auto& h : haps) out.add_hap(h, nullptr);
}

// synthetic.rs:199-289
void pd_region_generate(Batch& out, uint64_t seed, int64_t num_reads, int64_t read_len,
                        int64_t num_haps, int64_t hap_len) {
  Rng rng(seed);
  std::vector<uint8_t> reference(static_cast<size_t>(hap_len));
  for (auto& b : reference) b = rng.base();
  std::vector<std::vector<uint8_t>> hb, hf, realized;
  for (int64_t k = 0; k < num_haps; ++k) {
    std::vector<uint8_t> bases = reference;
    if (k > 0) {
      const uint64_t edits = rng.below(3);
      for (uint64_t e = 0; e < edits; ++e) {
        const uint64_t pos = rng.below(bases.size());
        switch (rng.below(3)) {
          case 0: bases[pos] = rng.base(); break;
          case 1: {
            const uint64_t n = 1 + rng.below(5);
            std::vector<uint8_t> ins(n);
            for (auto& b : ins) b = rng.base();
            bases.insert(bases.begin() + int64_t(pos), ins.begin(), ins.end());
            break;
          }
          default: {
            const uint64_t end = std::min<uint64_t>(pos + 1 + rng.below(5), bases.size());
            if (bases.size() - (end - pos) >= 2) bases.erase(bases.begin() + int64_t(pos), bases.begin() + int64_t(end));
            break;
          }
        }
      }
    }
    const size_t n = bases.size();
    std::vector<uint8_t> flags(n, 0), alt = bases;
    std::vector<char> deleted(n, 0);
    const uint64_t events = 1 + rng.below(4);
    for (uint64_t e = 0; e < events; ++e) {
      const uint64_t pos = rng.below(n);
      if (flags[pos] != 0) continue;
      if (rng.chance(0.5)) {
        uint8_t mask;
        for (;;) {
          mask = uint8_t(rng.below(16));
          if (mask != 0) break;
        }
        uint8_t flag = 1;  // SNP
        for (int bit = 0; bit < 4; ++bit)
          if (mask & (1 << bit)) flag |= uint8_t(8 << bit);  // ALT_A..ALT_T
        flags[pos] = flag;
        std::vector<uint8_t> choices;
        for (int bit = 0; bit < 4; ++bit)
          if (mask & (1 << bit)) choices.push_back(kBases[bit]);
        alt[pos] = choices[rng.below(choices.size())];
      } else {
        const uint64_t len = std::min<uint64_t>(1 + rng.below(8), n - pos);
        bool clash = false;
        for (uint64_t j = pos; j < pos + len; ++j) clash = clash || flags[j] != 0;
        if (clash) continue;
        flags[pos] |= 2;            // DEL_START
        flags[pos + len - 1] |= 4;  // DEL_END
        for (uint64_t j = pos; j < pos + len; ++j) deleted[j] = 1;
      }
    }
    std::vector<uint8_t> allele;
    for (size_t j = 0; j < n; ++j)
      if (!deleted[j]) allele.push_back(alt[j]);
    realized.push_back(allele.empty() ? bases : allele);
    hb.push_back(bases);
    hf.push_back(flags);
  }
  out.begin_region();
  for (int64_t i = 0; i < num_reads; ++i) {
    const uint64_t k = rng.below(uint64_t(num_haps));
    const auto& source = rng.chance(0.5) ? realized[k] : hb[k];
    out.add_read(sample_read(rng, source, read_len));
  }
  for (int64_t k = 0; k < num_haps; ++k) out.add_hap(hb[size_t(k)], &hf[size_t(k)]);
}

// ---- benchmark workloads ---------------------------------------------------------------------
namespace {

// Base quality: round(N(33, 6)) clipped to 2..41, 2% of bases set to 2.
uint8_t draw_qual(Rng& rng) {
  if (rng.chance(0.02)) return 2;
  const double q = std::nearbyint(33.0 + 6.0 * rng.normal());
  return uint8_t(std::min(41.0, std::max(2.0, q)));
}

// A haplotype carrying 1-3 SNPs or indels of 1-10 bp relative to `ref`.
std::vector<uint8_t> edit_haplotype(Rng& rng, const std::vector<uint8_t>& ref) {
  std::vector<uint8_t> hap = ref;
  const uint64_t edits = 1 + rng.below(3);
  for (uint64_t e = 0; e < edits; ++e) {
    const uint64_t pos = rng.below(hap.size());
    switch (rng.below(3)) {
      case 0: {
        const uint8_t old = hap[pos];
        uint8_t nb = rng.base();
        while (nb == old) nb = rng.base();
        hap[pos] = nb;
        break;
      }
      case 1: {
        const uint64_t n = 1 + rng.below(10);
        std::vector<uint8_t> ins(n);
        for (auto& b : ins) b = rng.base();
        hap.insert(hap.begin() + int64_t(pos), ins.begin(), ins.end());
        break;
      }
      default: {
        const uint64_t end = std::min<uint64_t>(pos + 1 + rng.below(10), hap.size());
        if (hap.size() - (end - pos) >= 1) hap.erase(hap.begin() + int64_t(pos), hap.begin() + int64_t(end));
        break;
      }
    }
  }
  return hap;
}

// A read of exactly n bases from `source` at a uniform offset with substitution errors drawn
// with probability 10^(-q/10), or of uniformly random bases when `foreign`.
Read draw_read(Rng& rng, const std::vector<uint8_t>& source, int64_t n, bool foreign, int qual_mode,
               uint8_t ins_q, uint8_t del_q) {
  Read r;
  const int64_t start = foreign ? 0 : int64_t(rng.below(uint64_t(int64_t(source.size()) - n + 1)));
  for (int64_t i = 0; i < n; ++i) {
    const uint8_t q = qual_mode == 0 ? draw_qual(rng) : uint8_t(rng.range(30, 41));
    uint8_t b = foreign ? rng.base() : source[size_t(start + i)];
    if (!foreign && rng.chance(std::pow(10.0, -double(q) / 10.0))) {
      uint8_t nb = rng.base();
      while (nb == b) nb = rng.base();
      b = nb;
    }
    r.bases.push_back(b);
    r.quals.push_back(q);
    r.ins.push_back(ins_q);
    r.del.push_back(del_q);
    r.gcp.push_back(10);
  }
  return r;
}

// Undetermined events at about 5% of positions: a SNP with one alternate base, or a deletion
// run of 1-8 bases; returns the flags and a realized allele taking every event.
void draw_pd_flags(Rng& rng, const std::vector<uint8_t>& bases, std::vector<uint8_t>& flags,
                   std::vector<uint8_t>& realized) {
  const size_t n = bases.size();
  flags.assign(n, 0);
  std::vector<uint8_t> alt = bases;
  std::vector<char> deleted(n, 0);
  size_t pos = 0;
  while (pos < n) {
    if (!rng.chance(0.05)) { ++pos; continue; }
    if (rng.chance(0.5)) {
      uint8_t a = rng.base();
      while (a == bases[pos]) a = rng.base();
      const int bit = a == 'A' ? 0 : a == 'C' ? 1 : a == 'G' ? 2 : 3;
      flags[pos] = uint8_t(1 | (8 << bit));
      alt[pos] = a;
      ++pos;
    } else {
      const size_t len = std::min<size_t>(1 + rng.below(8), n - pos);
      flags[pos] |= 2;
      flags[pos + len - 1] |= 4;
      for (size_t j = pos; j < pos + len; ++j) deleted[j] = 1;
      pos += len;
    }
  }
  realized.clear();
  for (size_t j = 0; j < n; ++j)
    if (!deleted[j]) realized.push_back(alt[j]);
  if (realized.empty()) realized = bases;
}

}  // namespace

void workload_generate(Batch& out, int config, uint64_t seed, int64_t n_regions) {
  Rng rng(seed);
  std::vector<uint8_t> flags, realized;
  std::vector<std::vector<uint8_t>> haps, fl, real;
  if (config == 0) n_regions = n_regions > 0 ? n_regions : 1;
  for (int64_t g = 0; g < n_regions; ++g) {
    int64_t R, H, m_ref;
    std::vector<int64_t> lens;
    const bool pd = config == 3;
    double foreign_frac = 0.0;
    int qual_mode = 0;
    bool per_read_gaps = false;
    if (config == 0) {
      R = 32; H = 8; m_ref = rng.range(300, 350);
      lens.assign(size_t(R), 151);
    } else if (config == 2) {
      R = 64; H = 8; m_ref = 600;
      lens.assign(size_t(R), 250);
      foreign_frac = 0.3;
      qual_mode = 1;
    } else {
      const bool half = config == 4;
      R = rng.range(1, half ? 200 : 400);
      H = rng.range(2, half ? 8 : 16);
      lens.resize(size_t(R));
      for (auto& l : lens) l = rng.range(50, 250);
      m_ref = *std::max_element(lens.begin(), lens.end()) + rng.range(50, 400);
      per_read_gaps = true;
    }
    std::vector<uint8_t> ref(static_cast<size_t>(m_ref));
    for (auto& b : ref) b = rng.base();
    haps.assign(1, ref);
    for (int64_t k = 1; k < H; ++k) haps.push_back(edit_haplotype(rng, ref));
    fl.clear();
    real.clear();
    if (pd) {
      for (const auto& h : haps) {
        draw_pd_flags(rng, h, flags, realized);
        fl.push_back(flags);
        real.push_back(realized);
      }
    }
    out.begin_region();
    for (int64_t i = 0; i < R; ++i) {
      const uint64_t k = rng.below(uint64_t(H));
      const bool foreign = foreign_frac > 0.0 && rng.chance(foreign_frac);
      const std::vector<uint8_t>* src = &haps[k];
      if (pd && rng.chance(0.5)) src = &real[k];
      int64_t n = lens[size_t(i)];
      n = std::min<int64_t>(n, int64_t(src->size()));
      uint8_t iq = 45, dq = 45;
      if (per_read_gaps) {
        iq = uint8_t(rng.range(10, 45));
        dq = uint8_t(rng.range(10, 45));
      }
      out.add_read(draw_read(rng, *src, n, foreign, qual_mode, iq, dq));
    }
    for (int64_t k = 0; k < H; ++k) out.add_hap(haps[size_t(k)], pd ? &fl[size_t(k)] : nullptr);
  }
  out.end_batch();
}

}  // namespace fgkl

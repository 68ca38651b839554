// scx_tune.h -- the search of the online launch-width autotuner (host code only).
//
// A persistent sweep launch runs on the cost model's workgroup count times 2^e. Per
// class of launch the table climbs from the class's start exponent: first one doubling;
// while a step saves more than kMargin of the time per sweep it keeps going that way,
// otherwise it tries the halvings from the start likewise; then it settles on the fastest
// exponent timed. A new class starts from the settled exponent of the nearest class of the
// same kind (bits 8-10 of the key equal; distance = difference of the contact bucket, bits
// 0-7, plus the difference of the interval bucket, bits 11 and up). Launches with fewer
// than kMinSweeps sweeps are not timed; kMaxMisses of them at the exponent being tried
// settle the class.
#pragma once

#include <cstdint>
#include <cstdlib>
#include <map>

namespace scx {

struct TuneEntry {
  static constexpr int kMinExp = -3, kMaxExp = 3, kExps = kMaxExp - kMinExp + 1;
  float us[kExps];                // least time per sweep seen at each exponent, microseconds (< 0: none)
  uint8_t misses[kExps];
  int8_t cur = 0;                 // exponent of the next launch
  int8_t start = 0;
  int8_t dir = 1;                 // +1 climbing up, -1 climbing down, 0 settled
  int model = 1, max_blocks = 1;  // of the last launch (bounds the useful exponents)
  uint32_t launches = 0;
  TuneEntry() {
    for (int i = 0; i < kExps; ++i) {
      us[i] = -1.f;
      misses[i] = 0;
    }
  }
  float at(int e) const { return us[e - kMinExp]; }
};

class TuneTable {
 public:
  static constexpr uint32_t kMinSweeps = 16;
  static constexpr int kMaxMisses = 3;
  static constexpr float kMargin = 0.97f;  // a step must save 3 % per sweep to count

  // Workgroups of exponent e around the model's pick, within [1, max_blocks].
  static int blocks(int model, int e, int max_blocks) {
    long long b = model;
    b = (e >= 0) ? (b << e) : (b >> -e);
    return (int)(b < 1 ? 1 : (b > max_blocks ? max_blocks : b));
  }

  // Exponent of the next launch of class `key` whose cost model picks `model` workgroups.
  int pick(uint32_t key, int model, int max_blocks) {
    auto it = table_.find(key);
    if (it == table_.end()) {
      TuneEntry E;
      int best_d = 1 << 30;
      for (const auto& kv : table_) {
        if (((kv.first >> 8) & 7u) != ((key >> 8) & 7u) || kv.second.dir != 0) continue;
        const int d = std::abs((int)(kv.first & 0xFFu) - (int)(key & 0xFFu)) +
                      std::abs((int)(kv.first >> 11) - (int)(key >> 11));
        if (d < best_d) {
          best_d = d;
          E.start = E.cur = kv.second.cur;
        }
      }
      it = table_.emplace(key, E).first;
    }
    TuneEntry& E = it->second;
    E.model = model;
    E.max_blocks = max_blocks;
    ++E.launches;
    return E.cur;
  }

  // A launch of class `key` at exponent e ran `sweeps` sweeps in `ms` milliseconds.
  void record(uint32_t key, int e, float ms, uint32_t sweeps) {
    auto it = table_.find(key);
    if (it == table_.end() || e < TuneEntry::kMinExp || e > TuneEntry::kMaxExp) return;
    TuneEntry& E = it->second;
    const int i = e - TuneEntry::kMinExp;
    if (sweeps < kMinSweeps) {
      if (E.dir != 0 && e == E.cur && ++E.misses[i] >= kMaxMisses) settle(E);
      return;
    }
    const float u = 1000.f * ms / (float)sweeps;
    E.us[i] = (E.us[i] < 0.f) ? u : (u < E.us[i] ? u : E.us[i]);
    if (E.dir != 0 && e == E.cur) climb(E);
  }

  const std::map<uint32_t, TuneEntry>& entries() const { return table_; }

 private:
  // e is worth a launch after `from`: in range and a different workgroup count.
  static bool useful(const TuneEntry& E, int e, int from) {
    return e >= TuneEntry::kMinExp && e <= TuneEntry::kMaxExp &&
           blocks(E.model, e, E.max_blocks) != blocks(E.model, from, E.max_blocks);
  }

  static void settle(TuneEntry& E) {
    int best = E.start;
    float bu = -1.f;
    for (int e = TuneEntry::kMinExp; e <= TuneEntry::kMaxExp; ++e) {
      const float u = E.at(e);
      if (u >= 0.f && (bu < 0.f || u < bu)) {
        bu = u;
        best = e;
      }
    }
    E.cur = (int8_t)best;
    E.dir = 0;
  }

  static void climb(TuneEntry& E) {
    const int x = E.cur;
    if (x == E.start) {
      if (useful(E, x + 1, x)) {
        E.cur = (int8_t)(x + 1);
        E.dir = 1;
      } else if (useful(E, x - 1, x)) {
        E.cur = (int8_t)(x - 1);
        E.dir = -1;
      } else {
        settle(E);
      }
      return;
    }
    const int prev = x - E.dir;
    if (E.at(prev) >= 0.f && E.at(x) < kMargin * E.at(prev)) {
      if (useful(E, x + E.dir, x))
        E.cur = (int8_t)(x + E.dir);
      else
        settle(E);
      return;
    }
    if (E.dir > 0 && prev == E.start && useful(E, E.start - 1, E.start) && E.at(E.start - 1) < 0.f) {
      E.cur = (int8_t)(E.start - 1);
      E.dir = -1;
      return;
    }
    settle(E);
  }

  std::map<uint32_t, TuneEntry> table_;
};

// Share (percent) of the contacts in the component phase of the sweeps, for the next step,
// from the last step's counters: launches whose components overflowed a workgroup's list,
// launches that ran the phase, and the largest count of one colour's contacts in a
// workgroup's list. On an overflow (the share reached the contact graph's giant component)
// the share drops to a tenth below itself, which also becomes a ceiling that drifts back up
// by 0.5% a step (about one retry in 20 steps while the graph stays the same); lowered by a
// fifth when a workgroup had more contacts of one colour than threads (`threads`); raised
// by a tenth, up to the ceiling, while every workgroup's widest colour stays within a
// quarter of its threads or no launch ran the phase (too few colours qualified). Results do
// not depend on the share, only the time of a sweep.
inline void comp_share_update(double& share, double& ceiling, uint32_t overflows, uint32_t launches,
                              uint32_t maxw, uint32_t threads) {
  ceiling = ceiling * 1.005 < 100.0 ? ceiling * 1.005 : 100.0;
  if (overflows > 0) {
    ceiling = 0.9 * share > 1.0 ? 0.9 * share : 1.0;
    share = ceiling;
  } else if (maxw > threads) {
    share = 0.8 * share > 1.0 ? 0.8 * share : 1.0;
  } else if (launches == 0 || maxw <= threads / 4) {
    const double up = 1.1 * share < ceiling ? 1.1 * share : ceiling;
    if (up > share) share = up;
  }
}

}  // namespace scx

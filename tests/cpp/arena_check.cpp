// Host-side check of the serving loop's arena (das_amd/csrc/arena.h): FIFO
// admission with at most 3 running requests, result blocks kept until the end
// (a caller that waits for every request before fetching any result), over
// every order in which running requests can finish.  No order may leave the
// queue head unable to fit while nothing runs, and releasing everything must
// coalesce back to one free block.  Prints "ok" and exits 0 on success.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "arena.h"

using tsw::Arena;

struct Rq { size_t heavy, keep; };

static bool run(const std::vector<Rq>& rq, size_t budget, const std::vector<int>& picks) {
  std::vector<char> mem(1);  // offsets only: base is never dereferenced
  Arena a;
  a.base = mem.data();
  a.size = Arena::up(budget);
  a.free_[0] = a.size;
  struct Run { int id; char* h; char* k; };
  std::vector<Run> running;
  std::vector<char*> kept;
  size_t next = 0, pick = 0;
  while (next < rq.size() || !running.empty()) {
    while (next < rq.size() && running.size() < 3) {
      char* h = a.alloc(rq[next].heavy);
      char* k = h ? a.alloc_high(rq[next].keep) : nullptr;
      if (!k) {
        a.release(h, rq[next].heavy);
        break;
      }
      running.push_back({(int)next, h, k});
      ++next;
    }
    if (running.empty()) return false;  // stall: head cannot fit, nothing to wait for
    const size_t j = (pick < picks.size() ? (size_t)picks[pick++] : 0) % running.size();
    a.release(running[j].h, rq[running[j].id].heavy);
    kept.push_back(running[j].k);  // results stay until every request is done
    running.erase(running.begin() + (long)j);
  }
  return true;
}

int main(int argc, char** argv) {
  // sizes: heavy working block and kept result block of the six requests of
  // the serving GPU test (acoustic and elastic, uneven grids and lengths)
  std::vector<Rq> rq;
  for (int i = 1; i + 1 < argc; i += 2)
    rq.push_back({(size_t)std::atoll(argv[i]), (size_t)std::atoll(argv[i + 1])});
  const size_t budget = (size_t)std::atoll(argv[argc - 1]);
  std::vector<int> picks(rq.size(), 0);
  long orders = 0;
  for (;;) {  // every sequence of picks in {0,1,2}^n
    if (!run(rq, budget, picks)) {
      std::printf("stall\n");
      return 1;
    }
    ++orders;
    size_t d = 0;
    while (d < picks.size() && ++picks[d] == 3) picks[d++] = 0;
    if (d == picks.size()) break;
  }
  // coalescing: allocate everything, release in a scrambled order, expect one block
  std::vector<char> mem(1);
  Arena a;
  a.base = mem.data();
  a.size = Arena::up(budget * 4);
  a.free_[0] = a.size;
  std::vector<std::pair<char*, size_t>> blocks;
  for (const auto& r : rq) {
    blocks.push_back({a.alloc(r.heavy), r.heavy});
    blocks.push_back({a.alloc_high(r.keep), r.keep});
  }
  for (size_t i = 0; i < blocks.size(); ++i) {
    const size_t j = (i * 7 + 3) % blocks.size();
    if (blocks[j].first) a.release(blocks[j].first, blocks[j].second);
    blocks[j].first = nullptr;
  }
  for (auto& b : blocks)
    if (b.first) a.release(b.first, b.second);
  if (a.free_.size() != 1 || a.free_.begin()->first != 0 || a.free_.begin()->second != a.size) {
    std::printf("not coalesced\n");
    return 1;
  }
  std::printf("ok %ld orders\n", orders);
  return 0;
}

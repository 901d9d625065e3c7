// Host emulation driver for the sift kernels of commit 97b620f (persistent kernel with
// deferral of traces whose knots outgrow the LDS tile, then a retry launch over a global
// tile).  Same input/output format as tests/native/sift_emu.cpp.
#include "emd_sift_dev.h"
#include <thread>
#include <cstring>
#include <limits>
thread_local double* emu_dyn_lds = nullptr;
namespace {
size_t dyn_lds(bool gt, int tile, int max_n) {
  const size_t nbcap = (size_t)(max_n + 63) / 64;
  return nbcap * emdk::kWorkBlockBytes + (gt ? 0 : (size_t)tile * emdk::kSlotBytes);
}
struct Input { int B, wg, tile, resident, persistent, max_extrema, nbsym, max_iteration, max_imfs;
  std::vector<int64_t> offsets; std::vector<double> values; std::vector<int> order; };
template <int WG, bool GT>
void run_wg(int bid, int grid, const emdk::Batch& b, const emdk::SiftParams& P) {
  emu::Group g(WG);
  std::vector<double> lds(dyn_lds(GT, P.tile_slots, P.max_n) / 8 + 1,
                          std::numeric_limits<double>::quiet_NaN());
  for (int t = 0; t < WG; ++t) {
    emu::Lane& L = g.lanes[t];
    L.tid.x = t; L.bid.x = bid; L.gdim.x = grid; L.bdim.x = WG; L.lane = t & 63; L.wave = t >> 6;
  }
  emu_dyn_lds = lds.data();
  emu::run_group(g, [&](emu::Lane&) { emdk::emd_persistent_kernel<WG, GT>(b, P); });
}
template <int WG, bool GT>
void launch(int grid, const emdk::Batch& b, const emdk::SiftParams& P) {
  std::vector<std::thread> th;
  for (int w = 0; w < grid; ++w) th.emplace_back([&, w] { run_wg<WG, GT>(w, grid, b, P); });
  for (auto& t : th) t.join();
}
template <class T> bool rd(FILE* f, T* p, size_t n) { return n == 0 || std::fread(p, sizeof(T), n, f) == n; }
}  // namespace
int main(int argc, char** argv) {
  if (argc != 3) return 2;
  FILE* f = std::fopen(argv[1], "rb");
  int32_t hdr[11]; int64_t nv = 0; Input in;
  if (!f || !rd(f, hdr, 11) || hdr[0] != 0x454d5531 || !rd(f, &nv, 1)) return 2;
  in.B = hdr[1]; in.wg = hdr[2]; in.tile = hdr[3]; in.resident = hdr[4]; in.persistent = hdr[5];
  in.max_extrema = hdr[6]; in.nbsym = hdr[7]; in.max_iteration = hdr[8]; in.max_imfs = hdr[9];
  in.offsets.resize(in.B + 1); in.values.resize(nv);
  if (!rd(f, in.offsets.data(), in.B + 1) || !rd(f, in.values.data(), nv)) return 2;
  if (hdr[10]) { in.order.resize(in.B); if (!rd(f, in.order.data(), in.B)) return 2; }
  std::fclose(f);
  int max_n = 0;
  for (int s = 0; s < in.B; ++s) max_n = std::max<int>(max_n, in.offsets[s + 1] - in.offsets[s]);
  const int cap = 2 * in.max_extrema + 4 * in.nbsym;
  emdk::SiftParams P{};
  P.max_extrema = in.max_extrema; P.nbsym = in.nbsym; P.max_iteration = in.max_iteration;
  P.max_imfs = in.max_imfs; P.tile_slots = in.tile > 0 ? std::min(in.tile, cap) : cap;
  P.max_n = std::max(max_n, 1); P.zero_fill = 1;
  const int K = in.max_imfs;
  std::vector<double> imfs((size_t)std::max<int64_t>(nv * K, 1), std::numeric_limits<double>::quiet_NaN());
  std::vector<int> n_imfs(in.B, -7), ovf(in.B, -7), dn(in.B * K, -7), de(in.B * K, -7),
      dz(in.B * K, -7), stats(in.B * 4, -7), defer_list(in.B, -7);
  unsigned counters[4] = {0, 0, 0, 0};
  std::vector<int64_t> oo(in.B);
  for (int q = 0; q < in.B; ++q) oo[q] = in.offsets[q] * K;  // the Python default then

  emdk::Batch b{};
  b.values = in.values.data(); b.offsets = in.offsets.data(); b.N = 0; b.B = in.B; b.base = 0;
  b.order = in.order.empty() ? nullptr : in.order.data(); b.imfs = imfs.data();
  b.out_offsets = oo.data(); b.n_imfs = n_imfs.data(); b.overflow = ovf.data();
  b.dbg_n = dn.data(); b.dbg_e = de.data(); b.dbg_z = dz.data(); b.stats = stats.data();
  b.spill = nullptr; b.spill_stride = 0; b.queue = &counters[0]; b.count_ptr = nullptr;
  b.defer_count = &counters[1]; b.defer_list = defer_list.data(); b.retry = 0;
  const int grid = std::max(1, std::min(in.B, in.resident));
  std::fprintf(stderr, "emu 97b620f: B=%d wg=%d tile=%d grid=%d max_n=%d\n", in.B, in.wg, P.tile_slots, grid, max_n);
  switch (in.wg) {
    case 64: launch<64, false>(grid, b, P); break;
    case 128: launch<128, false>(grid, b, P); break;
    case 256: launch<256, false>(grid, b, P); break;
    default: return 2;
  }
  std::fprintf(stderr, "emu 97b620f: main launch done, %u deferred\n", counters[1]);
  // retry launch: global tile holding any trace, queue over the deferred list
  emdk::SiftParams R = P; R.tile_slots = cap;
  const int64_t stride = ((int64_t)cap * emdk::kSlotBytes + 255) / 256 * 256;
  const int rgrid = std::max(1, in.resident);
  std::vector<char> spill((size_t)stride * rgrid);
  std::memset(spill.data(), 0xA5, spill.size());
  emdk::Batch r = b;
  r.order = defer_list.data(); r.queue = &counters[2]; r.count_ptr = &counters[1];
  r.defer_count = &counters[3]; r.defer_list = defer_list.data(); r.retry = 1;
  r.spill = spill.data(); r.spill_stride = stride;
  switch (in.wg) {
    case 64: launch<64, true>(rgrid, r, R); break;
    case 128: launch<128, true>(rgrid, r, R); break;
    case 256: launch<256, true>(rgrid, r, R); break;
  }
  std::fprintf(stderr, "emu 97b620f: retry launch done, %u deferred again\n", counters[3]);
  FILE* o = std::fopen(argv[2], "wb");
  std::fwrite(n_imfs.data(), 4, in.B, o); std::fwrite(ovf.data(), 4, in.B, o);
  std::fwrite(dn.data(), 4, in.B * K, o); std::fwrite(de.data(), 4, in.B * K, o);
  std::fwrite(dz.data(), 4, in.B * K, o); std::fwrite(stats.data(), 4, in.B * 4, o);
  std::fwrite(imfs.data(), 8, nv * K, o); std::fclose(o);
  return 0;
}

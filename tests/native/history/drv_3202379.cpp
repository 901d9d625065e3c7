// Host emulation driver for the sift kernels of commit 3202379 (register-resident
// samples, LDS knot tile, deferral and retry launch).  Same input/output format as tests/native/sift_emu.cpp.
#include "emd_sift_dev.h"
#include <thread>
#include <cstring>
#include <limits>
thread_local double* emu_dyn_lds = nullptr;
namespace {
size_t dyn_lds(bool gt, int tile, int) { return gt ? 0 : (size_t)tile * emdk::kSlotBytes; }
int chunk_for(int n, int wg) { int per = (n + wg - 1) / wg; int c = 1; while (c < per) c <<= 1; return c; }
struct Input { int B, wg, tile, resident, persistent, max_extrema, nbsym, max_iteration, max_imfs;
  std::vector<int64_t> offsets; std::vector<double> values; std::vector<int> order; };
template <int WG, int C, bool GT, bool PERS>
void run_wg(int bid, int grid, const emdk::Batch& b, const emdk::SiftParams& P, int max_n) {
  emu::Group g(WG);
  std::vector<double> lds(dyn_lds(GT, P.tile_slots, max_n) / 8 + 1,
                          std::numeric_limits<double>::quiet_NaN());
  for (int t = 0; t < WG; ++t) {
    emu::Lane& L = g.lanes[t];
    L.tid.x = t; L.bid.x = bid; L.gdim.x = grid; L.bdim.x = WG; L.lane = t & 63; L.wave = t >> 6;
  }
  emu_dyn_lds = lds.data();
  emu::run_group(g, [&](emu::Lane&) {
    if (PERS) emdk::emd_persistent_kernel<WG, C, GT>(b, P);
    else emdk::emd_plain_kernel<WG, C, GT>(b, P);
  });
}
template <int WG, int C, bool GT, bool PERS>
void launch_c(int grid, const emdk::Batch& b, const emdk::SiftParams& P, int max_n, int conc) {
  if (PERS) {
    std::vector<std::thread> th;
    for (int w = 0; w < grid; ++w) th.emplace_back([&, w] { run_wg<WG, C, GT, PERS>(w, grid, b, P, max_n); });
    for (auto& t : th) t.join();
    return;
  }
  for (int first = 0; first < grid; first += conc) {
    std::vector<std::thread> th;
    for (int w = first; w < std::min(grid, first + conc); ++w)
      th.emplace_back([&, w] { run_wg<WG, C, GT, PERS>(w, grid, b, P, max_n); });
    for (auto& t : th) t.join();
  }
}
template <int WG, bool GT, bool PERS>
void launch(int grid, const emdk::Batch& b, const emdk::SiftParams& P, int max_n, int conc) {
  switch (chunk_for(max_n, WG)) {
    case 1: launch_c<WG, 1, GT, PERS>(grid, b, P, max_n, conc); break;
    case 2: launch_c<WG, 2, GT, PERS>(grid, b, P, max_n, conc); break;
    case 4: launch_c<WG, 4, GT, PERS>(grid, b, P, max_n, conc); break;
    case 8: launch_c<WG, 8, GT, PERS>(grid, b, P, max_n, conc); break;
    case 16: launch_c<WG, 16, GT, PERS>(grid, b, P, max_n, conc); break;
    case 32: launch_c<WG, 32, GT, PERS>(grid, b, P, max_n, conc); break;
    default: std::fprintf(stderr, "chunk %d not emulated\n", chunk_for(max_n, WG)); std::exit(2);
  }
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
  const int cap = 2 * std::min(in.max_extrema, std::max(max_n, 1)) + 4 * in.nbsym;
  emdk::SiftParams P{};
  P.max_extrema = in.max_extrema; P.nbsym = in.nbsym; P.max_iteration = in.max_iteration;
  P.max_imfs = in.max_imfs; P.tile_slots = in.tile > 0 ? std::min(in.tile, cap) : cap;
  P.zero_fill = 1;
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
  std::fprintf(stderr, "emu 3202379: B=%d wg=%d tile=%d grid=%d max_n=%d\n", in.B, in.wg, P.tile_slots, grid, max_n);
  const int grid_main = in.persistent ? grid : in.B;
  switch (in.wg) {
    case 64: if (in.persistent) launch<64, false, true>(grid_main, b, P, max_n, grid); else launch<64, false, false>(grid_main, b, P, max_n, grid); break;
    case 128: if (in.persistent) launch<128, false, true>(grid_main, b, P, max_n, grid); else launch<128, false, false>(grid_main, b, P, max_n, grid); break;
    case 256: if (in.persistent) launch<256, false, true>(grid_main, b, P, max_n, grid); else launch<256, false, false>(grid_main, b, P, max_n, grid); break;
    default: return 2;
  }
  std::fprintf(stderr, "emu 3202379: main launch done, %u deferred\n", counters[1]);
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
  // the retry shape: the capacity tile in LDS (it fits for these traces), same workgroup size
  r.spill = nullptr; r.spill_stride = 0;
  switch (in.wg) {
    case 64: launch<64, false, true>(rgrid, r, R, max_n, rgrid); break;
    case 128: launch<128, false, true>(rgrid, r, R, max_n, rgrid); break;
    case 256: launch<256, false, true>(rgrid, r, R, max_n, rgrid); break;
  }
  std::fprintf(stderr, "emu 3202379: retry launch done, %u deferred again\n", counters[3]);
  FILE* o = std::fopen(argv[2], "wb");
  std::fwrite(n_imfs.data(), 4, in.B, o); std::fwrite(ovf.data(), 4, in.B, o);
  std::fwrite(dn.data(), 4, in.B * K, o); std::fwrite(de.data(), 4, in.B * K, o);
  std::fwrite(dz.data(), 4, in.B * K, o); std::fwrite(stats.data(), 4, in.B * 4, o);
  std::fwrite(imfs.data(), 8, nv * K, o); std::fclose(o);
  return 0;
}

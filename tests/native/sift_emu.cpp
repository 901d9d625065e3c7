// Host emulation of the sift kernels: the source of emd-cuda_amd/csrc/emd_sift_dev.h,
// built with g++ against the host stand-in of the HIP API (tests/native/emu/hip/
// hip_runtime.h), with EMD_BOUNDS_CHECK (every knot-tile and work-area index checked
// against its own array) and, in the test build, AddressSanitizer and UndefinedBehavior-
// Sanitizer.  Every lane is a fiber of its workgroup's OS thread; workgroups of a
// persistent launch run on concurrent threads and share the queue counter, as on the GPU.
//
// Test infrastructure: tests/test_sift_emu.py writes a batch, runs this program on it and
// compares the outputs with the oracle.  It runs the launch the C ABI would plan for the
// same arguments (tile clamped to the capacity rule and to what 160 KB of LDS holds, a
// global knot tile of the full capacity when the LDS tile is smaller),
// with the number of concurrently resident workgroups given on the command line.
//
// usage: sift_emu <input file> <output file>
// input (little-endian): int32 magic 0x454d5531, B, wg, tile_slots, resident, persistent,
//   max_extrema, nbsym, max_iteration, max_imfs, has_order, has_times; int64 n_values;
//   int64 offsets[B+1]; double values[n_values]; int32 order[B] if has_order;
//   double times[n_values] if has_times (each trace's normalised time axis, laid out like
//   values: the time-axis kernels run)
// output: int32 n_imfs[B], overflow[B], debug_n/extno/nzm [B*max_imfs], stats[B*4];
//   double imfs[n_values*max_imfs] (row k of trace s at offsets[s]*max_imfs + k*len(s))
// exit status: 0 done, 2 bad input, 3 a barrier not reached by every lane (printed),
//   and abort / sanitizer exit codes on an out-of-bounds access.
#include "emd_sift_dev.h"

#include <thread>

namespace {

int64_t region_bytes(int slots, int slot_bytes) {
  return ((int64_t)slots * slot_bytes + 255) / 256 * 256;
}

size_t dyn_lds(int tile_slots, int max_n, int slot_bytes) {
  const size_t nbcap = (size_t)(max_n + 63) / 64;
  return nbcap * emdk::kWorkBlockBytes + (size_t)tile_slots * slot_bytes;
}

struct Input {
  int B, wg, tile, resident, persistent, max_extrema, nbsym, max_iteration, max_imfs;
  std::vector<int64_t> offsets;
  std::vector<double> values;
  std::vector<int> order;
  std::vector<double> times;
};

template <int WG, class AX>
void run_workgroup(int bid, int grid, const emdk::Batch& b, const emdk::SiftParams& P,
                   bool persistent) {
  emu::Group g(WG);
  // LDS starts as a NaN pattern each launch: a read of an element no lane wrote shows up
  // as a NaN in the outputs
  std::vector<double> lds(dyn_lds(P.tile_slots, P.max_n, AX::kSlot) / 8 + 1,
                          std::numeric_limits<double>::quiet_NaN());
  auto* sh = new emdk::SiftShared<WG>;
  std::memset((void*)sh, 0xA5, sizeof(*sh));
  for (int t = 0; t < WG; ++t) {
    emu::Lane& L = g.lanes[t];
    L.tid.x = t;
    L.bid.x = bid;
    L.gdim.x = grid;
    L.bdim.x = WG;
    L.lane = t & 63;
    L.wave = t >> 6;
  }
  emu::run_group(g, [&](emu::Lane&) {
    if (persistent)
      emdk::persistent_body<WG, AX>(b, P, *sh, lds.data());
    else
      emdk::plain_body<WG, AX>(b, P, *sh, lds.data());
  });
  delete sh;
}

template <int WG, class AX>
void run(const Input& in, const emdk::Batch& b0, const emdk::SiftParams& P, int64_t stride) {
  emdk::Batch b = b0;
  if (in.persistent) {
    const int grid = std::max(1, std::min(in.B, in.resident));
    std::vector<char> spill((size_t)std::max<int64_t>(stride * grid, 1));
    std::memset(spill.data(), 0xA5, spill.size());
    b.spill = stride ? spill.data() : nullptr;
    b.spill_stride = stride;
    unsigned queue = 0;
    b.queue = &queue;
    b.B = in.B;
    b.base = 0;
    std::vector<std::thread> wgs;
    for (int w = 0; w < grid; ++w)
      wgs.emplace_back([&, w] { run_workgroup<WG, AX>(w, grid, b, P, true); });
    for (auto& th : wgs) th.join();
    return;
  }
  // plain: one workgroup per queue position, `resident` of them at a time (each with its
  // own global region, as on the GPU where blockIdx picks the region)
  const int conc = std::max(1, in.resident);
  std::vector<char> spill((size_t)std::max<int64_t>(stride * in.B, 1));
  std::memset(spill.data(), 0xA5, spill.size());
  b.spill = stride ? spill.data() : nullptr;
  b.spill_stride = stride;
  b.B = in.B;
  b.base = 0;
  for (int first = 0; first < in.B; first += conc) {
    std::vector<std::thread> wgs;
    for (int w = first; w < std::min(in.B, first + conc); ++w)
      wgs.emplace_back([&, w] { run_workgroup<WG, AX>(w, in.B, b, P, false); });
    for (auto& th : wgs) th.join();
  }
}

template <class T>
bool rd(FILE* f, T* p, size_t n) {
  return n == 0 || std::fread(p, sizeof(T), n, f) == n;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: sift_emu <input> <output>\n");
    return 2;
  }
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int32_t hdr[12];
  int64_t n_values = 0;
  Input in;
  if (!rd(f, hdr, 12) || hdr[0] != 0x454d5531 || !rd(f, &n_values, 1)) return 2;
  in.B = hdr[1];
  in.wg = hdr[2];
  in.tile = hdr[3];
  in.resident = hdr[4];
  in.persistent = hdr[5];
  in.max_extrema = hdr[6];
  in.nbsym = hdr[7];
  in.max_iteration = hdr[8];
  in.max_imfs = hdr[9];
  in.offsets.resize(in.B + 1);
  in.values.resize(n_values);
  if (!rd(f, in.offsets.data(), in.B + 1) || !rd(f, in.values.data(), n_values)) return 2;
  if (hdr[10]) {
    in.order.resize(in.B);
    if (!rd(f, in.order.data(), in.B)) return 2;
  }
  const bool timed = hdr[11] != 0;
  if (timed) {
    in.times.resize(n_values);
    if (!rd(f, in.times.data(), n_values)) return 2;
  }
  std::fclose(f);
  if (in.B < 0 || in.offsets[0] < 0 || in.offsets[in.B] > n_values) return 2;
  for (int s = 0; s < in.B; ++s)
    if (in.offsets[s + 1] < in.offsets[s]) return 2;
  for (int q : in.order)
    if (q < 0 || q >= in.B) return 2;

  int max_n = 0;
  for (int s = 0; s < in.B; ++s) max_n = std::max<int>(max_n, in.offsets[s + 1] - in.offsets[s]);
  // the C ABI's plan (emd_capi.hip make_shape): capacity rule, tile clamped to it and to
  // the LDS next to the work area, a global knot tile of the full capacity behind a
  // smaller tile
  const int cap = 2 * std::min(in.max_extrema, std::max(max_n, 1)) + 4 * in.nbsym;
  int tile = in.tile > 0 ? std::min(in.tile, cap) : cap;
  const long lds_max = 163840;
  const int slot_bytes = timed ? emdk::kSlotBytesTime : emdk::kSlotBytes;
  const long fit = (lds_max - (long)(in.wg == 64    ? sizeof(emdk::SiftShared<64>)
                                     : in.wg == 128 ? sizeof(emdk::SiftShared<128>)
                                                    : sizeof(emdk::SiftShared<256>)) -
                    (long)dyn_lds(0, max_n, slot_bytes)) /
                   slot_bytes;
  tile = std::max(0, std::min<int>(tile, (int)std::max(0L, fit)));
  emdk::SiftParams P;
  P.max_extrema = in.max_extrema;
  P.nbsym = in.nbsym;
  P.max_iteration = in.max_iteration;
  P.max_imfs = in.max_imfs;
  P.tile_slots = tile;
  P.spill_slots = tile < cap ? cap : 0;
  P.max_n = std::max(max_n, 1);
  P.zero_fill = 0;
  const int64_t stride = P.spill_slots ? region_bytes(P.spill_slots, slot_bytes) : 0;

  const int K = in.max_imfs;
  std::vector<double> imfs((size_t)std::max<int64_t>(n_values * K, 1),
                           std::numeric_limits<double>::quiet_NaN());
  std::vector<int> n_imfs(in.B, -7), overflow(in.B, -7), dn(in.B * K, -7), de(in.B * K, -7),
      dz(in.B * K, -7), stats(in.B * 4, -7);
  emdk::Batch b{};
  b.values = in.values.data();
  b.offsets = in.offsets.data();
  b.N = 0;
  b.B = in.B;
  b.base = 0;
  b.order = in.order.empty() ? nullptr : in.order.data();
  b.total = in.B;
  b.imfs = imfs.data();
  b.out_offsets = nullptr;
  b.n_imfs = n_imfs.data();
  b.overflow = overflow.data();
  b.dbg_n = dn.data();
  b.dbg_e = de.data();
  b.dbg_z = dz.data();
  b.stats = stats.data();
  b.times = timed ? in.times.data() : nullptr;
  b.times_shared = 0;
  std::fprintf(stderr, "sift_emu: B=%d wg=%d tile=%d (cap %d) spill_slots=%d "
               "resident=%d persistent=%d max_n=%d\n", in.B, in.wg, tile, cap, P.spill_slots,
               in.resident, in.persistent, max_n);
  switch (in.wg * 2 + (timed ? 1 : 0)) {
    case 128: run<64, emdk::IndexAxis>(in, b, P, stride); break;
    case 256: run<128, emdk::IndexAxis>(in, b, P, stride); break;
    case 512: run<256, emdk::IndexAxis>(in, b, P, stride); break;
    case 129: run<64, emdk::TimeAxis>(in, b, P, stride); break;
    case 257: run<128, emdk::TimeAxis>(in, b, P, stride); break;
    case 513: run<256, emdk::TimeAxis>(in, b, P, stride); break;
    default: std::fprintf(stderr, "workgroup size %d not emulated\n", in.wg); return 2;
  }
  FILE* o = std::fopen(argv[2], "wb");
  if (!o) return 2;
  std::fwrite(n_imfs.data(), 4, in.B, o);
  std::fwrite(overflow.data(), 4, in.B, o);
  std::fwrite(dn.data(), 4, in.B * K, o);
  std::fwrite(de.data(), 4, in.B * K, o);
  std::fwrite(dz.data(), 4, in.B * K, o);
  std::fwrite(stats.data(), 4, in.B * 4, o);
  std::fwrite(imfs.data(), 8, n_values * K, o);
  std::fclose(o);
  return 0;
}

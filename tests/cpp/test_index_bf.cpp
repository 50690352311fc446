// C++ drop-in test of curplsh::IndexBF / searchL2Distance / computeL2Norm,
// written like the reference's drivers (test/test_search.cpp,
// test/test_norm.cpp) but asserting. Needs a GPU. Exit status 0 = pass.
//
// Checks:
//  1. test_norm fixture (test/test_norm.cpp:15-22): all-ones rows -> norm = dim.
//  2. test_search fixture (test/test_search.cpp:17-36) through the internal
//     searchL2Distance with relative distances: top-10 [1,0,2,...,9].
//  3. IndexBF with host pointers vs the canonical CPU oracle (bit-exact I, D).
//  4. IndexBF with device pointers gives the same results; repeated add
//     equals one add; save + load into a fresh object gives the same
//     results; reset empties the index.
//  5. DatasetSIFT over files written with helper::saveXvecsData (base,
//     query, FP64 ground truth): loaded into device memory, searched from
//     device pointers, evaluate() >= 0.999 (FP32 vs FP64 truth) and the
//     arrays left as they were.
//  6. IndexRPLSH (LSH front end) vs the oracle's LSH restatement with the
//     index's projections and offsets: bit-exact I and D; reset empties it.
//  7. The internal selection functions: blockSelect (smallest / largest,
//     implicit and given labels), l2Select + broadcastAlongRowsSum on a small
//     fixture with exact ties (lower label first).
//  8. DeviceMemory (include/device_memory.h:12-89) over the HBM budget:
//     reservations used as tensor storage through narrow()ed views, blocks
//     reused after release, the budget respected.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <unistd.h>

#include "curplsh/dataset.h"
#include "curplsh/device_memory.h"
#include "curplsh/index_bf.h"
#include "curplsh/index_rplsh.h"
#include "curplsh/internal/broadcast.h"
#include "curplsh/internal/norm.h"
#include "curplsh/internal/select.h"
#include "curplsh/internal/search.h"
#include "../../oracle/oracle.h"

#define CHECK(c)                                                          \
  do {                                                                    \
    if (!(c)) {                                                           \
      std::fprintf(stderr, "FAIL %s at %s:%d\n", #c, __FILE__, __LINE__); \
      return 1;                                                           \
    }                                                                     \
  } while (0)

using namespace curplsh;

static int test_norm(DeviceResources& res) {
  for (int dim : {32, 64, 96, 100, 128, 256, 1500}) {
    const int num = 1001;
    std::vector<float> h((size_t)num * dim, 1.0f);
    DeviceTensor<float, 2> v({num, dim});
    DeviceTensor<float, 1> n({num});
    Stream s = res.getDefaultStream(0);
    v.copyFrom(h.data(), s);
    computeL2Norm(v, n, true, s);
    std::vector<float> out(num);
    n.copyTo(out.data(), s);
    streamSynchronize(s);
    for (int i = 0; i < num; ++i) CHECK(out[i] == (float)dim);
  }
  std::printf("norm fixture ok\n");
  return 0;
}

static int test_search_fixture(DeviceResources& res) {
  const int dim = 128, nb = 1000, nq = 100, k = 10;
  std::vector<float> b((size_t)nb * dim), q((size_t)nq * dim, 1.0f);
  for (int i = 0; i < nb; ++i)
    for (int d = 0; d < dim; ++d) b[(size_t)i * dim + d] = std::fmin((float)i, 10.0f);
  Stream s = res.getDefaultStream(0);
  DeviceTensor<float, 2> bases({nb, dim}), queries({nq, dim}), dist({nq, k});
  DeviceTensor<int, 2> idx({nq, k});
  bases.copyFrom(b.data(), s);
  queries.copyFrom(q.data(), s);
  searchL2Distance(&res, bases, nullptr, queries, k, idx, dist, false);
  std::vector<int> I((size_t)nq * k);
  std::vector<float> D((size_t)nq * k);
  idx.copyTo(I.data(), s);
  dist.copyTo(D.data(), s);
  streamSynchronize(s);
  const int eI[10] = {1, 0, 2, 3, 4, 5, 6, 7, 8, 9};
  const float eD[10] = {-128, 0, 0, 384, 1024, 1920, 3072, 4480, 6144, 8064};
  for (int r = 0; r < nq; ++r)
    for (int j = 0; j < k; ++j) {
      CHECK(I[(size_t)r * k + j] == eI[j]);
      CHECK(D[(size_t)r * k + j] == eD[j]);
    }
  std::printf("search fixture ok\n");
  return 0;
}

static int test_index(DeviceResources& res) {
  const int dim = 96, n = 50000, nq = 300, k = 20;
  std::vector<float> x((size_t)n * dim), q((size_t)nq * dim);
  CHECK(rplsh_gen_uniform(x.data(), n, dim, 11, -1.0f, 1.0f) == RPLSH_OK);
  CHECK(rplsh_gen_uniform(q.data(), nq, dim, 12, -1.0f, 1.0f) == RPLSH_OK);
  std::vector<int32_t> Io((size_t)nq * k);
  std::vector<float> Do((size_t)nq * k);
  oracle_search_canonical(x.data(), n, dim, nullptr, q.data(), nq, k, 1, Io.data(), Do.data(), 0);

  IndexBF index(&res, dim);
  index.train(n, x.data());
  index.add(n / 2, x.data());
  index.add(n - n / 2, x.data() + (size_t)(n / 2) * dim);
  CHECK(index.getNumData() == n);
  std::vector<int> I((size_t)nq * k);
  std::vector<float> D((size_t)nq * k);
  index.search(nq, q.data(), k, I.data(), D.data());
  CHECK(std::memcmp(I.data(), Io.data(), I.size() * 4) == 0);
  CHECK(std::memcmp(D.data(), Do.data(), D.size() * 4) == 0);

  // Device-resident queries and outputs.
  DeviceTensor<float, 2> dq({nq, dim}), dd({nq, k});
  DeviceTensor<int, 2> di({nq, k});
  Stream s = res.getDefaultStream(0);
  dq.copyFrom(q.data(), s);
  streamSynchronize(s);
  index.search(nq, dq.data(), k, di.data(), dd.data());
  std::vector<int> I2((size_t)nq * k);
  std::vector<float> D2((size_t)nq * k);
  di.copyTo(I2.data(), s);
  dd.copyTo(D2.data(), s);
  streamSynchronize(s);
  CHECK(I2 == I);
  CHECK(std::memcmp(D2.data(), D.data(), D.size() * 4) == 0);

  // Index file round trip into a fresh object.
  char path[] = "/tmp/rplsh_indexXXXXXX";
  const int fd = mkstemp(path);
  CHECK(fd >= 0);
  close(fd);
  index.save(path);
  IndexBF other(&res, dim);
  other.load(path);
  std::remove(path);
  CHECK(other.getNumData() == n);
  std::vector<int> I3((size_t)nq * k);
  std::vector<float> D3((size_t)nq * k);
  other.search(nq, q.data(), k, I3.data(), D3.data());
  CHECK(I3 == I);
  CHECK(std::memcmp(D3.data(), D.data(), D.size() * 4) == 0);

  index.reset();
  CHECK(index.getNumData() == 0);
  index.search(nq, q.data(), k, I.data(), D.data());
  for (size_t i = 0; i < I.size(); ++i) CHECK(I[i] == -1);
  std::printf("IndexBF ok\n");
  return 0;
}

static int test_dataset(DeviceResources& res) {
  const int dim = 128, n = 30000, nq = 100, k = 100;
  std::vector<float> x((size_t)n * dim), q((size_t)nq * dim);
  CHECK(rplsh_gen_uniform(x.data(), n, dim, 21, 0.0f, 1.0f) == RPLSH_OK);
  CHECK(rplsh_gen_uniform(q.data(), nq, dim, 22, 0.0f, 1.0f) == RPLSH_OK);
  std::vector<int32_t> gt((size_t)nq * k);
  std::vector<double> gd((size_t)nq * k);
  oracle_search_fp64(x.data(), n, dim, q.data(), nq, k, gt.data(), gd.data(), 0);
  char dir[] = "/tmp/rplsh_dsXXXXXX";
  CHECK(mkdtemp(dir) != nullptr);
  const std::string d(dir);
  helper::saveXvecsData((d + "/sift_base.fvecs").c_str(), x.data(), n, dim);
  helper::saveXvecsData((d + "/sift_query.fvecs").c_str(), q.data(), nq, dim);
  helper::saveXvecsData((d + "/sift_groundtruth.ivecs").c_str(), gt.data(), nq, k);
  int rc = 0;
  {
    DatasetSIFT ds(dir, MemorySpace::Device);
    CHECK(ds.getNumBase() == n && ds.getNumQuery() == nq && ds.getDimension() == dim &&
          ds.getGroundTruthK() == k);
    CHECK(getDeviceForAddress(ds.getBase()) == 0);
    IndexBF index(&res, dim);
    index.add(n, ds.getBase());
    DeviceTensor<int, 2> di({nq, k});
    DeviceTensor<float, 2> dd({nq, k});
    index.search(nq, ds.getQuery(), k, di.data(), dd.data());
    std::vector<int> before((size_t)nq * k), after((size_t)nq * k), diff(nq);
    Stream s = res.getDefaultStream(0);
    di.copyTo(before.data(), s);
    streamSynchronize(s);
    const float recall = ds.evaluate(di.data(), diff.data());
    di.copyTo(after.data(), s);
    streamSynchronize(s);
    // FP32 vs FP64 ground truth: a boundary near-tie may differ, nothing more.
    int missed = 0;
    for (int i = 0; i < nq; ++i) missed += diff[i];
    if (!(recall >= 0.999f && before == after && missed == (int)std::lround((1.0 - recall) * nq * k))) rc = 1;
    std::printf("DatasetSIFT recall@%d = %.4f\n", k, recall);
  }
  for (const char* f : {"/sift_base.fvecs", "/sift_query.fvecs", "/sift_groundtruth.ivecs"})
    std::remove((d + f).c_str());
  rmdir(dir);
  CHECK(rc == 0);
  std::printf("Dataset ok\n");
  return 0;
}

static int test_internal_select(DeviceResources& res) {
  // rows: {5, 1, 3, 1, 9} and {2, 2, 8, 0.5, 2}; labels 10..14 / 20..24
  const float vals[10] = {5, 1, 3, 1, 9, 2, 2, 8, 0.5f, 2};
  const int labs[10] = {10, 11, 12, 13, 14, 24, 23, 22, 21, 20};
  Stream s = res.getDefaultStream(0);
  DeviceTensor<float, 2> in({2, 5});
  DeviceTensor<int, 2> lab({2, 5});
  in.copyFrom(vals, s);
  lab.copyFrom(labs, s);
  DeviceTensor<float, 2> ok({2, 3});
  DeviceTensor<int, 2> ov({2, 3});
  float k3[6];
  int v3[6];
  blockSelect(in, ok, ov, 3, false, s);  // column labels
  ok.copyTo(k3, s);
  ov.copyTo(v3, s);
  streamSynchronize(s);
  CHECK(k3[0] == 1 && v3[0] == 1 && k3[1] == 1 && v3[1] == 3 && k3[2] == 3 && v3[2] == 2);
  CHECK(k3[3] == 0.5f && v3[3] == 3 && k3[4] == 2 && v3[4] == 0 && k3[5] == 2 && v3[5] == 1);
  blockSelect(in, lab, ok, ov, 3, true, s);  // given labels, largest first
  ok.copyTo(k3, s);
  ov.copyTo(v3, s);
  streamSynchronize(s);
  CHECK(k3[0] == 9 && v3[0] == 14 && k3[1] == 5 && v3[1] == 10 && k3[2] == 3 && v3[2] == 12);
  CHECK(k3[3] == 8 && v3[3] == 22 && k3[4] == 2 && v3[4] == 20 && k3[5] == 2 && v3[5] == 23);
  // l2Select: norms + products, then + 10 per row
  const float norms[5] = {0, 1, 0, 2, 0};
  DeviceTensor<float, 1> dn({5});
  dn.copyFrom(norms, s);
  DeviceTensor<float, 1> add({2});
  const float adds[2] = {10, 20};
  add.copyFrom(adds, s);
  l2Select(in, dn, ok, ov, 3, s);
  broadcastAlongRowsSum(add, ok, s);
  ok.copyTo(k3, s);
  ov.copyTo(v3, s);
  streamSynchronize(s);
  // row 0 keys {5, 2, 3, 3, 9} -> (2,1) (3,2) (3,3); row 1 {2, 3, 8, 2.5, 2} -> (2,0) (2,4) (2.5,3)
  CHECK(k3[0] == 12 && v3[0] == 1 && k3[1] == 13 && v3[1] == 2 && k3[2] == 13 && v3[2] == 3);
  CHECK(k3[3] == 22 && v3[3] == 0 && k3[4] == 22 && v3[4] == 4 && k3[5] == 22.5f && v3[5] == 3);
  std::printf("internal select ok\n");
  return 0;
}

static int test_rplsh(DeviceResources& res) {
  const int n = 20000, dim = 48, nq = 64, k = 10;
  std::vector<float> x((size_t)n * dim), q((size_t)nq * dim);
  CHECK(rplsh_gen_mixture(x.data(), n, dim, 16, 0.3f, 5, 6, 0.0) == RPLSH_OK);
  CHECK(rplsh_gen_mixture(q.data(), nq, dim, 16, 0.3f, 5, 7, 0.0) == RPLSH_OK);
  IndexRPLSHConfig cfg;
  cfg.numTables = 8;
  cfg.numBits = 12;
  cfg.numProbes = 2;
  cfg.maxCandidates = 1024;
  IndexRPLSH index(&res, dim, cfg);
  index.add(n, x.data());
  CHECK(index.getNumData() == n);
  std::vector<float> w((size_t)cfg.numTables * cfg.numBits * dim), off((size_t)cfg.numTables * cfg.numBits);
  CHECK(rplsh_lsh_projections(index.handle(), w.data()) == RPLSH_OK);
  CHECK(rplsh_lsh_offsets(index.handle(), off.data()) == RPLSH_OK);
  std::vector<int> I((size_t)nq * k), Io((size_t)nq * k);
  std::vector<float> D((size_t)nq * k), Do((size_t)nq * k);
  index.search(nq, q.data(), k, I.data(), D.data());
  int64_t raw = 0, distinct = 0;
  oracle_lsh_search(x.data(), n, dim, w.data(), off.data(), cfg.numTables, cfg.numBits, cfg.numProbes,
                    cfg.maxCandidates, 0, q.data(), nq, k, Io.data(), Do.data(), &raw, &distinct);
  CHECK(I == Io);
  CHECK(std::memcmp(D.data(), Do.data(), D.size() * sizeof(float)) == 0);
  const rplsh_lsh_stats st = index.lastStats();
  CHECK(st.raw_candidates == raw && st.distinct_candidates == distinct);
  index.reset();
  CHECK(index.getNumData() == 0);
  index.search(nq, q.data(), k, I.data(), D.data());
  for (int v : I) CHECK(v == -1);
  std::printf("IndexRPLSH ok (%.1f distinct candidates per query)\n", (double)distinct / nq);
  return 0;
}

static int test_device_memory(DeviceResources& res) {
  const int n = 3000, dim = 64;
  Stream s = res.getDefaultStream(0);
  BudgetDeviceMemory mem(0, (size_t)64 << 20);
  CHECK(mem.getDevice() == 0);
  std::vector<float> h((size_t)n * dim), out(n);
  for (size_t i = 0; i < h.size(); ++i) h[i] = (float)((i * 7) % 13) - 6.0f;
  void* first = nullptr;
  for (int round = 0; round < 2; ++round) {
    DeviceMemoryReservation rx = mem.getMemory(s, h.size() * sizeof(float));
    DeviceMemoryReservation rn = mem.getMemory(s, (size_t)n * sizeof(float));
    CHECK(rx.get() && rn.get() && rx.size() == h.size() * sizeof(float));
    if (round == 0) first = rx.get();
    else CHECK(rx.get() == first);  // the released block came back
    DeviceTensor<float, 2> x(static_cast<float*>(rx.get()), {n, dim}, MemorySpace::Device);
    DeviceTensor<float, 1> nr(static_cast<float*>(rn.get()), {n}, MemorySpace::Device);
    CHECK(!x.owner());
    x.copyFrom(h.data(), s);
    // norms of the second half only, through narrowed views of both tensors
    Tensor<float, 2> xh = x.narrow(0, n / 2, n - n / 2);
    Tensor<float, 1> nh = nr.narrow(0, n / 2, n - n / 2);
    CHECK(xh.data() == x.data() + (size_t)(n / 2) * dim && xh.getSize(0) == n - n / 2);
    nr.zero(s);
    computeL2Norm(xh, nh, true, s);
    nr.copyTo(out.data(), s);
    streamSynchronize(s);
    for (int i = 0; i < n; ++i) {
      float e = 0.0f;
      if (i >= n / 2)
        for (int j = 0; j < dim; ++j) e = std::fma(h[(size_t)i * dim + j], h[(size_t)i * dim + j], e);
      CHECK(out[i] == e);
    }
  }
  const rplsh_memory_stats st = mem.stats();
  CHECK(st.mallocs == 2 && st.reservations == 4 && st.reserved_bytes == 0);
  CHECK(st.high_water_malloc_bytes >= (size_t)n * dim * 4 && mem.getSizeAvailable() >= (size_t)n * dim * 4);
  std::printf("DeviceMemory ok: %s\n", mem.toString().c_str());
  return 0;
}

int main() {
  if (getNumDevices() < 1) {
    std::fprintf(stderr, "no GPU\n");
    return 2;
  }
  DeviceScope scope(0);
  DeviceResources res(0);
  if (test_norm(res)) return 1;
  if (test_search_fixture(res)) return 1;
  if (test_index(res)) return 1;
  if (test_dataset(res)) return 1;
  if (test_rplsh(res)) return 1;
  if (test_internal_select(res)) return 1;
  if (test_device_memory(res)) return 1;
  std::printf("all C++ tests passed\n");
  return 0;
}

// Operator-level tests of the drop-in physical operators, in the shape of the
// reference's operator tests (fpdb-executor/ATTIC/test/{JoinTest,GroupTest,
// BloomJoinTest}.cpp): build a plan of POps, execute it, check the output.
//
//   pop_test unit                 built-in vectors and error paths
//   pop_test starjoin N OUT.csv   predicate-transfer star join + group-by over
//                                 N seeded fact rows; result rows to OUT.csv
//                                 (checked against numpy by tests/test_pop_gpu.py)
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <set>
#include <string>
#include <tuple>
#include <vector>

#include "BloomFilterPOps.h"
#include "Cache.h"
#include "Globals.h"
#include "Group.h"
#include "HashJoin.h"
#include "HashJoinBuildProbe.h"
#include "Plumbing.h"
#include "PredTrans.h"
#include "Shuffle.h"

using namespace fpdb::amd;

static int g_failures = 0;
#define CHECK(cond)                                                                 \
  do {                                                                              \
    if (!(cond)) {                                                                  \
      std::printf("  CHECK failed at line %d: %s\n", __LINE__, #cond);               \
      ++g_failures;                                                                 \
    }                                                                               \
  } while (0)

static HostColumn hostInt64(const std::vector<int64_t>& v) {
  HostColumn h;
  h.type = FPDB_TYPE_INT64;
  h.length = (int64_t)v.size();
  h.bytes.resize(v.size() * 8);
  std::memcpy(h.bytes.data(), v.data(), h.bytes.size());
  return h;
}

static HostColumn hostInt32(const std::vector<int32_t>& v) {
  HostColumn h;
  h.type = FPDB_TYPE_INT32;
  h.length = (int64_t)v.size();
  h.bytes.resize(v.size() * 4);
  std::memcpy(h.bytes.data(), v.data(), h.bytes.size());
  return h;
}

static std::shared_ptr<TupleSet> table3(const std::string& prefix, const std::vector<std::vector<int64_t>>& rows) {
  Schema s;
  std::vector<std::vector<int64_t>> cols(3);
  for (auto& r : rows)
    for (int c = 0; c < 3; ++c) cols[c].push_back(r[c]);
  std::vector<HostColumn> hc;
  for (int c = 0; c < 3; ++c) {
    s.fields.push_back(Field{prefix + (char)('a' + c), FPDB_TYPE_INT64, false});
    hc.push_back(hostInt64(cols[c]));
  }
  return TupleSet::fromHost(defaultContext(), s, hc);
}

static std::vector<std::vector<int64_t>> rowsOf(const std::shared_ptr<TupleSet>& ts) {
  auto h = ts->toHost();
  std::vector<std::vector<int64_t>> rows(ts->numRows());
  for (int64_t r = 0; r < ts->numRows(); ++r)
    for (auto& c : h) rows[r].push_back(c.isValid(r) ? c.int64At(r) : INT64_MIN);
  return rows;
}

// ATTIC data/join/{a,b}.csv and data/group/a.csv
static const std::vector<std::vector<int64_t>> kA = {{10, 13, 16}, {10, 14, 17}, {12, 15, 18}};
static const std::vector<std::vector<int64_t>> kB = {{10, 23, 26}, {11, 24, 27}, {10, 25, 28}};

static std::shared_ptr<TupleSet> runJoin(JoinType jt, bool bloomFirst) {
  Executor ex;
  auto a = std::make_shared<SourcePOp>("scan_a", std::vector<std::shared_ptr<TupleSet>>{table3("a", kA)});
  auto b = std::make_shared<SourcePOp>("scan_b", std::vector<std::shared_ptr<TupleSet>>{table3("b", kB)});
  auto join = std::make_shared<HashJoinArrowPOp>("join", std::vector<std::string>{"aa", "ab", "ac", "ba", "bb", "bc"}, 0,
                                                 HashJoinPredicate{{"ba"}, {"aa"}}, jt);
  auto collate = std::make_shared<CollatePOp>("collate");
  ex.put(a);
  ex.put(b);
  ex.put(join);
  ex.put(collate);
  if (bloomFirst) {
    // BloomJoinTest: a filter on b.ba applied to a.aa before the join
    auto bfc = std::make_shared<BloomFilterCreatePOp>("bf_create", std::vector<std::string>{}, 0,
                                                      std::vector<std::string>{"ba"});
    auto bfu = std::make_shared<BloomFilterUsePOp>("bf_use", std::vector<std::string>{}, 0,
                                                   std::vector<std::string>{"aa"});
    ex.put(bfc);
    ex.put(bfu);
    connect(b, bfc);
    bfc->addBloomFilterUsePOp(bfu);
    bfu->consume(bfc);
    connect(a, bfu);
    bfc->produce(join);
    join->addBuildProducer(bfc);
    bfu->produce(join);
    join->addProbeProducer(bfu);
  } else {
    b->produce(join);
    join->addBuildProducer(b);
    a->produce(join);
    join->addProbeProducer(a);
  }
  connect(join, collate);
  auto r = ex.execute();
  if (!r) {
    std::printf("  execute failed: %s\n", r.error().c_str());
    ++g_failures;
    return nullptr;
  }
  return collate->tupleSet().value();
}

static void testJoin() {
  // JoinTest.cpp:82-123: 4 rows, probe columns (a) first
  auto out = runJoin(INNER, false);
  if (!out) return;
  CHECK(out->numRows() == 4 && out->numColumns() == 6);
  CHECK(out->schema().fields[0].name == "aa" && out->schema().fields[3].name == "ba");
  auto rows = rowsOf(out);
  std::sort(rows.begin(), rows.end());
  std::vector<std::vector<int64_t>> want = {{10, 13, 16, 10, 23, 26}, {10, 13, 16, 10, 25, 28},
                                            {10, 14, 17, 10, 23, 26}, {10, 14, 17, 10, 25, 28}};
  CHECK(rows == want);
}

static void testOuterSemiJoins() {
  auto right = runJoin(RIGHT, false);    // probe side (a) kept: a row aa=12 with null b
  if (right) {
    auto rows = rowsOf(right);
    CHECK(right->numRows() == 5);
    int nulls = 0;
    for (auto& r : rows) nulls += r[3] == INT64_MIN;
    CHECK(nulls == 1);
  }
  auto left = runJoin(LEFT, false);      // build side (b) kept: b row ba=11 with null a
  if (left) {
    auto rows = rowsOf(left);
    CHECK(left->numRows() == 5);
    int found = 0;
    for (auto& r : rows) found += (r[0] == INT64_MIN && r[3] == 11);
    CHECK(found == 1);
  }
  auto full = runJoin(FULL, false);
  if (full) CHECK(full->numRows() == 6);
  auto lsemi = runJoin(LEFT_SEMI, false);   // build rows with a match: ba = 10 twice
  if (lsemi) {
    CHECK(lsemi->numRows() == 2 && lsemi->numColumns() == 3);
    for (auto& r : rowsOf(lsemi)) CHECK(r[0] == 10);
  }
  auto rsemi = runJoin(RIGHT_SEMI, false);  // probe rows with a match: aa = 10 twice
  if (rsemi) {
    CHECK(rsemi->numRows() == 2 && rsemi->schema().fields[0].name == "aa");
  }
}

// HashJoinBuildPOp -> TupleSetIndexMessage -> HashJoinProbePOp.  With
// `probeLate` the probe rows pass one extra hop, so they reach the probe
// operator after the build indexes; otherwise they arrive first.
static std::shared_ptr<TupleSet> runBuildProbeJoin(JoinType jt, const std::vector<std::shared_ptr<TupleSet>>& build,
                                                   const std::vector<std::shared_ptr<TupleSet>>& probe,
                                                   std::vector<std::string> project, bool probeLate) {
  Executor ex;
  auto b = std::make_shared<SourcePOp>("scan_b", build);
  auto a = std::make_shared<SourcePOp>("scan_a", probe);
  auto hb = std::make_shared<HashJoinBuildPOp>("build", std::vector<std::string>{}, 0, std::vector<std::string>{"ba"});
  auto hp = std::make_shared<HashJoinProbePOp>("probe", project, 0, HashJoinPredicate{{"ba"}, {"aa"}}, jt);
  auto collate = std::make_shared<CollatePOp>("collate");
  ex.put(b);
  ex.put(a);
  ex.put(hb);
  ex.put(hp);
  ex.put(collate);
  connect(b, hb);
  connect(hb, hp);
  if (probeLate) {
    auto hop = std::make_shared<FilterPOp>("hop", std::vector<std::string>{}, 0, "aa", FPDB_CMP_GE, INT64_MIN);
    ex.put(hop);
    connect(a, hop);
    connect(hop, hp);
  } else {
    connect(a, hp);
  }
  connect(hp, collate);
  auto r = ex.execute();
  if (!r) {
    std::printf("  execute failed: %s\n", r.error().c_str());
    ++g_failures;
    return nullptr;
  }
  return collate->tupleSet().value();
}

static void testBuildProbeJoin() {
  // JoinTest.cpp:82-123 through the two-operator join: build (b) columns first
  std::vector<std::string> all = {"aa", "ab", "ac", "ba", "bb", "bc"};
  auto out = runBuildProbeJoin(INNER, {table3("b", kB)}, {table3("a", kA)}, all, false);
  if (out) {
    CHECK(out->numRows() == 4 && out->numColumns() == 6);
    CHECK(out->schema().fields[0].name == "ba" && out->schema().fields[3].name == "aa");
    auto rows = rowsOf(out);
    std::sort(rows.begin(), rows.end());
    std::vector<std::vector<int64_t>> want = {{10, 23, 26, 10, 13, 16}, {10, 23, 26, 10, 14, 17},
                                              {10, 25, 28, 10, 13, 16}, {10, 25, 28, 10, 14, 17}};
    CHECK(rows == want);
  }
  auto left = runBuildProbeJoin(LEFT, {table3("b", kB)}, {table3("a", kA)}, all, false);
  if (left) {
    int found = 0;
    for (auto& r : rowsOf(left)) found += (r[0] == 11 && r[3] == INT64_MIN);
    CHECK(left->numRows() == 5 && found == 1);
  }
  auto right = runBuildProbeJoin(RIGHT, {table3("b", kB)}, {table3("a", kA)}, all, false);
  if (right) {
    int found = 0;
    for (auto& r : rowsOf(right)) found += (r[0] == INT64_MIN && r[3] == 12);
    CHECK(right->numRows() == 5 && found == 1);
  }
  auto full = runBuildProbeJoin(FULL, {table3("b", kB)}, {table3("a", kA)}, all, false);
  if (full) CHECK(full->numRows() == 6);
  auto semi = runBuildProbeJoin(LEFT_SEMI, {table3("b", kB)}, {table3("a", kA)}, {"ba", "bb", "bc"}, false);
  if (semi) {
    CHECK(semi->numRows() == 2 && semi->numColumns() == 3);
    for (auto& r : rowsOf(semi)) CHECK(r[0] == 10);
  }
  // an empty probe side: the output schema still arrives, with no rows
  auto none = runBuildProbeJoin(INNER, {table3("b", kB)}, {table3("a", {})}, all, false);
  if (none) CHECK(none->numRows() == 0 && none->numColumns() == 6 && none->schema().fields[0].name == "ba");

  // Several build indexes (DefaultBufferSize rows each) and probe batches with
  // duplicate keys, both arrival orders, every join type: the rows as a
  // multiset equal a nested-loop join.
  int64_t savedBuffer = DefaultBufferSize;
  DefaultBufferSize = 7;
  uint64_t s = 12345;
  auto rnd = [&s](int64_t m) {
    s = s * 6364136223846793005ULL + 1442695040888963407ULL;
    return (int64_t)((s >> 33) % (uint64_t)m);
  };
  std::vector<std::vector<int64_t>> brows, prows;
  std::vector<std::shared_ptr<TupleSet>> bsets, psets;
  for (int t = 0; t < 5; ++t) {
    std::vector<std::vector<int64_t>> part;
    for (int i = 0; i < 5; ++i) part.push_back({rnd(12), (int64_t)brows.size() + (int64_t)part.size(), rnd(100)});
    brows.insert(brows.end(), part.begin(), part.end());
    bsets.push_back(table3("b", part));
  }
  for (int t = 0; t < 4; ++t) {
    std::vector<std::vector<int64_t>> part;
    for (int i = 0; i < 6; ++i) part.push_back({rnd(15), 1000 + (int64_t)prows.size() + (int64_t)part.size(), rnd(100)});
    prows.insert(prows.end(), part.begin(), part.end());
    psets.push_back(table3("a", part));
  }
  const int64_t N = INT64_MIN;
  for (JoinType jt : {INNER, LEFT, RIGHT, FULL, LEFT_SEMI}) {
    std::vector<std::vector<int64_t>> want;
    std::vector<bool> pm(prows.size(), false);
    for (auto& br : brows) {
      bool matched = false;
      for (size_t j = 0; j < prows.size(); ++j) {
        if (br[0] != prows[j][0]) continue;
        matched = pm[j] = true;
        if (jt != LEFT_SEMI) want.push_back({br[0], br[1], br[2], prows[j][0], prows[j][1], prows[j][2]});
      }
      if (jt == LEFT_SEMI && matched) want.push_back(br);
      if ((jt == LEFT || jt == FULL) && !matched) want.push_back({br[0], br[1], br[2], N, N, N});
    }
    if (jt == RIGHT || jt == FULL)
      for (size_t j = 0; j < prows.size(); ++j)
        if (!pm[j]) want.push_back({N, N, N, prows[j][0], prows[j][1], prows[j][2]});
    std::sort(want.begin(), want.end());
    for (bool late : {false, true}) {
      auto got = runBuildProbeJoin(jt, bsets, psets, jt == LEFT_SEMI ? std::vector<std::string>{"ba", "bb", "bc"} : all,
                                   late);
      if (!got) continue;
      auto rows = rowsOf(got);
      std::sort(rows.begin(), rows.end());
      if (rows != want) std::printf("  join type %d, probe late %d: %zu rows, want %zu\n", (int)jt, (int)late,
                                    rows.size(), want.size());
      CHECK(rows == want);
    }
  }
  DefaultBufferSize = savedBuffer;
}

// ---- utf8 columns and keys the join kernels do not take directly ----------------

// A host table of int64 / utf8 columns; null cells are "NULL", strings "s:..."
struct HTable {
  Schema schema;
  std::vector<std::vector<std::string>> cells;   // [column][row]; "NULL" = null
  int64_t rows() const { return cells.empty() ? 0 : (int64_t)cells[0].size(); }
  const std::string& cell(const std::string& name, int64_t r) const {
    return cells[schema.getFieldIndex(name)][r];
  }
  std::shared_ptr<TupleSet> device(int64_t lo, int64_t hi) const {
    std::vector<HostColumn> hc;
    for (size_t c = 0; c < schema.fields.size(); ++c) {
      HostColumn h;
      if (schema.fields[c].type == FPDB_TYPE_UTF8) {
        std::vector<std::string> v;
        for (int64_t r = lo; r < hi; ++r) v.push_back(cells[c][r] == "NULL" ? "" : cells[c][r].substr(2));
        h = HostColumn::utf8(v);
      } else {
        std::vector<int64_t> v;
        for (int64_t r = lo; r < hi; ++r) v.push_back(cells[c][r] == "NULL" ? 0 : std::stoll(cells[c][r]));
        h = hostInt64(v);
      }
      bool anyNull = false;
      for (int64_t r = lo; r < hi; ++r) anyNull = anyNull || cells[c][r] == "NULL";
      if (anyNull)
        for (int64_t r = lo; r < hi; ++r) h.valid.push_back(cells[c][r] != "NULL");
      hc.push_back(h);
    }
    return TupleSet::fromHost(defaultContext(), schema, hc);
  }
  std::vector<std::shared_ptr<TupleSet>> batches(int64_t size) const {
    std::vector<std::shared_ptr<TupleSet>> out;
    for (int64_t lo = 0; lo < rows(); lo += size) out.push_back(device(lo, std::min(rows(), lo + size)));
    return out;
  }
};

static std::vector<std::vector<std::string>> cellsOf(const std::shared_ptr<TupleSet>& ts) {
  auto h = ts->toHost();
  std::vector<std::vector<std::string>> rows(ts->numRows());
  for (int64_t r = 0; r < ts->numRows(); ++r)
    for (auto& c : h)
      rows[r].push_back(!c.isValid(r) ? "NULL"
                                      : c.type == FPDB_TYPE_UTF8 ? "s:" + c.stringAt(r) : std::to_string(c.int64At(r)));
  return rows;
}

static uint64_t g_rng = 88172645463325252ULL;
static int64_t rnd(int64_t m) {
  g_rng ^= g_rng << 13;
  g_rng ^= g_rng >> 7;
  g_rng ^= g_rng << 17;
  return (int64_t)(g_rng % (uint64_t)m);
}

// Probe table a(aname utf8 null, ak1, ak2, ak3, acity utf8) and build table
// b(bname utf8 null, bk1, bk2, bk3, bregion utf8); wide keys repeat so joins match.
static HTable makeJoinTable(const std::string& p, int64_t n, int names) {
  HTable t;
  t.schema.fields = {Field{p + "name", FPDB_TYPE_UTF8, true}, Field{p + "k1", FPDB_TYPE_INT64, false},
                     Field{p + "k2", FPDB_TYPE_INT64, false}, Field{p + "k3", FPDB_TYPE_INT64, false},
                     Field{p + "txt", FPDB_TYPE_UTF8, false}};
  t.cells.resize(5);
  for (int64_t r = 0; r < n; ++r) {
    int64_t x = rnd(names);
    t.cells[0].push_back(rnd(10) == 0 ? "NULL" : x == 0 ? "s:" : "s:n\xc3\xa9" + std::to_string(x));
    t.cells[1].push_back(std::to_string((int64_t)(rnd(3)) * 4000000000LL));   // > 32 bits: not packable in pairs
    t.cells[2].push_back(std::to_string(-rnd(2)));
    t.cells[3].push_back(std::to_string(rnd(2)));
    t.cells[4].push_back("s:" + p + std::to_string(r) + std::string((size_t)rnd(5), 'x'));
  }
  return t;
}

// Expected join rows by nested loop: pairs (probe row, build row), -1 for a
// missing side, rendered in the output schema's column order.
static std::vector<std::vector<std::string>> nestedLoopJoin(const HTable& a, const HTable& b, JoinType jt,
                                                            const std::vector<std::string>& akeys,
                                                            const std::vector<std::string>& bkeys,
                                                            const Schema& out) {
  std::vector<std::pair<int64_t, int64_t>> pairs;
  std::vector<bool> am(a.rows(), false), bm(b.rows(), false);
  for (int64_t i = 0; i < a.rows(); ++i)
    for (int64_t j = 0; j < b.rows(); ++j) {
      bool eq = true;
      for (size_t k = 0; k < akeys.size() && eq; ++k) {
        const std::string &x = a.cell(akeys[k], i), &y = b.cell(bkeys[k], j);
        eq = x != "NULL" && x == y;
      }
      if (!eq) continue;
      am[i] = bm[j] = true;
      if (jt == INNER || jt == LEFT || jt == RIGHT || jt == FULL) pairs.push_back({i, j});
    }
  for (int64_t j = 0; j < b.rows(); ++j)
    if ((jt == LEFT || jt == FULL) ? !bm[j] : (jt == LEFT_SEMI && bm[j])) pairs.push_back({-1, j});
  for (int64_t i = 0; i < a.rows(); ++i)
    if ((jt == RIGHT || jt == FULL) ? !am[i] : (jt == RIGHT_SEMI && am[i])) pairs.push_back({i, -1});
  std::vector<std::vector<std::string>> rows;
  for (auto [i, j] : pairs) {
    std::vector<std::string> row;
    for (auto& f : out.fields) {
      bool probeSide = a.schema.getFieldIndex(f.name) >= 0;
      int64_t r = probeSide ? i : j;
      row.push_back(r < 0 ? "NULL" : (probeSide ? a : b).cell(f.name, r));
    }
    rows.push_back(row);
  }
  std::sort(rows.begin(), rows.end());
  return rows;
}

static std::shared_ptr<TupleSet> runGenericJoin(bool legacy, JoinType jt, const HTable& a, const HTable& b,
                                                const std::vector<std::string>& akeys,
                                                const std::vector<std::string>& bkeys,
                                                const std::vector<std::string>& project) {
  Executor ex;
  auto sa = std::make_shared<SourcePOp>("scan_a", a.batches(40));
  auto sb = std::make_shared<SourcePOp>("scan_b", b.batches(35));
  auto collate = std::make_shared<CollatePOp>("collate");
  ex.put(sa);
  ex.put(sb);
  ex.put(collate);
  HashJoinPredicate pred{bkeys, akeys};
  if (legacy) {
    auto hb = std::make_shared<HashJoinBuildPOp>("build", std::vector<std::string>{}, 0, bkeys);
    auto hp = std::make_shared<HashJoinProbePOp>("probe", project, 0, pred, jt);
    ex.put(hb);
    ex.put(hp);
    connect(sb, hb);
    connect(hb, hp);
    connect(sa, hp);
    connect(hp, collate);
  } else {
    auto join = std::make_shared<HashJoinArrowPOp>("join", project, 0, pred, jt);
    ex.put(join);
    sb->produce(join);
    join->addBuildProducer(sb);
    sa->produce(join);
    join->addProbeProducer(sa);
    connect(join, collate);
  }
  auto r = ex.execute();
  if (!r) {
    std::printf("  execute failed: %s\n", r.error().c_str());
    ++g_failures;
    return nullptr;
  }
  return collate->tupleSet().value();
}

static void testUtf8AndWideKeyJoins() {
  HTable a = makeJoinTable("a", 150, 12), b = makeJoinTable("b", 120, 15);
  std::vector<std::string> acols = {"aname", "ak1", "ak2", "ak3", "atxt"}, bcols = {"bname", "bk1", "bk2", "bk3", "btxt"};
  struct KeyCase {
    std::vector<std::string> akeys, bkeys;
  };
  std::vector<KeyCase> cases = {{{"aname"}, {"bname"}},                       // one utf8 key
                                {{"ak1", "ak2", "ak3"}, {"bk1", "bk2", "bk3"}},  // three integer keys
                                {{"ak1", "ak2"}, {"bk1", "bk2"}},                // two 8-byte keys
                                {{"aname", "ak3"}, {"bname", "bk3"}}};          // utf8 + integer
  for (bool legacy : {false, true}) {
    for (size_t kc = 0; kc < cases.size(); ++kc) {
      for (JoinType jt : {INNER, LEFT, RIGHT, FULL, LEFT_SEMI, RIGHT_SEMI}) {
        if (legacy && jt == RIGHT_SEMI) continue;
        std::vector<std::string> project;
        if (jt != LEFT_SEMI) project.insert(project.end(), acols.begin(), acols.end());
        if (jt != RIGHT_SEMI) project.insert(project.end(), bcols.begin(), bcols.end());
        auto got = runGenericJoin(legacy, jt, a, b, cases[kc].akeys, cases[kc].bkeys, project);
        if (!got) continue;
        auto rows = cellsOf(got);
        std::sort(rows.begin(), rows.end());
        auto want = nestedLoopJoin(a, b, jt, cases[kc].akeys, cases[kc].bkeys, got->schema());
        if (rows != want)
          std::printf("  legacy %d keys %zu join %d: %zu rows, want %zu\n", (int)legacy, kc, (int)jt, rows.size(),
                      want.size());
        CHECK(rows == want);
        // output column order: probe first (HashJoinArrowPOp), build first (HashJoinProbePOp)
        if (jt != LEFT_SEMI && jt != RIGHT_SEMI)
          CHECK(got->schema().fields[0].name == (legacy ? "bname" : "aname"));
      }
    }
  }
}

static void testUtf8GroupFilterShuffle() {
  // group by (utf8 with nulls, int with nulls) over three batches: groups in
  // first-appearance order, null a group of its own
  HTable g;
  g.schema.fields = {Field{"g1", FPDB_TYPE_UTF8, true}, Field{"g2", FPDB_TYPE_INT64, true},
                     Field{"v", FPDB_TYPE_INT64, false}};
  g.cells.resize(3);
  const int64_t n = 3000;
  for (int64_t r = 0; r < n; ++r) {
    int64_t x = rnd(9);
    g.cells[0].push_back(x == 0 ? "NULL" : "s:grp\xe2\x82\xac" + std::to_string(x));
    g.cells[1].push_back(rnd(7) == 0 ? "NULL" : std::to_string(rnd(3)));
    g.cells[2].push_back(std::to_string(rnd(1000)));
  }
  std::vector<std::pair<std::string, std::string>> order;
  std::map<std::pair<std::string, std::string>, std::pair<int64_t, int64_t>> agg;
  for (int64_t r = 0; r < n; ++r) {
    auto k = std::make_pair(g.cells[0][r], g.cells[1][r]);
    if (!agg.count(k)) order.push_back(k);
    agg[k].first += std::stoll(g.cells[2][r]);
    agg[k].second += 1;
  }
  {
    Executor ex;
    auto src = std::make_shared<SourcePOp>("scan", g.batches(1000));
    std::vector<std::shared_ptr<AggregateFunction>> fns = {std::make_shared<Sum>("s", "v"),
                                                          std::make_shared<Count>("c")};
    auto group = std::make_shared<GroupPOp>("group", std::vector<std::string>{"g1", "g2", "s", "c"}, 0,
                                            std::vector<std::string>{"g1", "g2"}, fns);
    auto collate = std::make_shared<CollatePOp>("collate");
    ex.put(src);
    ex.put(group);
    ex.put(collate);
    connect(src, group);
    connect(group, collate);
    auto r = ex.execute();
    if (!r) std::printf("  execute failed: %s\n", r.error().c_str());
    CHECK(r.has_value());
    if (r) {
      auto rows = cellsOf(collate->tupleSet().value());
      CHECK(rows.size() == order.size());
      for (size_t i = 0; i < rows.size() && i < order.size(); ++i) {
        auto& w = agg[order[i]];
        CHECK(rows[i][0] == order[i].first && rows[i][1] == order[i].second);
        CHECK(rows[i][2] == std::to_string(w.first) && rows[i][3] == std::to_string(w.second));
      }
    }
  }
  // three int64 keys whose ranges need more than 64 bits together: the group
  // operator moves from packed keys to dictionary ids mid-stream
  {
    HTable w;
    w.schema.fields = {Field{"x", FPDB_TYPE_INT64, false}, Field{"y", FPDB_TYPE_INT64, false},
                       Field{"z", FPDB_TYPE_INT64, false}};
    w.cells.resize(3);
    for (int64_t r = 0; r < 2000; ++r) {
      const int64_t big = r < 1000 ? 1 : (int64_t)1 << 40;   // the second half widens the ranges
      w.cells[0].push_back(std::to_string(rnd(3) * big));
      w.cells[1].push_back(std::to_string(-rnd(3) * big));
      w.cells[2].push_back(std::to_string(rnd(2) * big));
    }
    std::vector<std::vector<std::string>> ord;
    std::map<std::vector<std::string>, int64_t> cnt;
    for (int64_t r = 0; r < 2000; ++r) {
      std::vector<std::string> k = {w.cells[0][r], w.cells[1][r], w.cells[2][r]};
      if (!cnt.count(k)) ord.push_back(k);
      ++cnt[k];
    }
    Executor ex;
    auto src = std::make_shared<SourcePOp>("scan", w.batches(1000));
    std::vector<std::shared_ptr<AggregateFunction>> fns = {std::make_shared<Count>("c")};
    auto group = std::make_shared<GroupPOp>("group", std::vector<std::string>{"x", "y", "z", "c"}, 0,
                                            std::vector<std::string>{"x", "y", "z"}, fns);
    auto collate = std::make_shared<CollatePOp>("collate");
    ex.put(src);
    ex.put(group);
    ex.put(collate);
    connect(src, group);
    connect(group, collate);
    auto r = ex.execute();
    if (!r) std::printf("  execute failed: %s\n", r.error().c_str());
    CHECK(r.has_value());
    if (r) {
      auto rows = cellsOf(collate->tupleSet().value());
      CHECK(rows.size() == ord.size());
      for (size_t i = 0; i < rows.size() && i < ord.size(); ++i) {
        CHECK(std::vector<std::string>(rows[i].begin(), rows[i].begin() + 3) == ord[i]);
        CHECK(rows[i][3] == std::to_string(cnt[ord[i]]));
      }
    }
  }
  // a filter on a utf8 column, and a shuffle carrying utf8 columns
  {
    auto ts = g.device(0, n);
    std::string c = "grp\xe2\x82\xac" "3";
    auto f = FilterPOp::filter(ts, "g1", FPDB_CMP_GE, 0, 0.0, &c);
    CHECK(f.has_value());
    if (f) {
      auto rows = cellsOf(*f);
      std::vector<std::vector<std::string>> want;
      auto all = cellsOf(ts);
      for (auto& row : all)
        if (row[0] != "NULL" && row[0].substr(2) >= c) want.push_back(row);
      CHECK(rows == want);
    }
    CHECK(!FilterPOp::filter(ts, "g1", FPDB_CMP_EQ, 3, 0.0).has_value());   // a number against utf8
    auto slots = ShuffleKernel2::shuffle({"v"}, 3, ts);
    CHECK(slots.has_value());
    if (slots) {
      std::vector<std::vector<std::string>> got, want = cellsOf(ts);
      for (auto& sl : *slots)
        for (auto& row : cellsOf(sl)) got.push_back(row);
      std::sort(got.begin(), got.end());
      std::sort(want.begin(), want.end());
      CHECK(got == want);
    }
    CHECK(!ShuffleKernel2::shuffle({"g1"}, 3, ts).has_value());   // utf8 shuffle keys are refused
  }
  // a bloom filter applied to a batch with utf8 columns keeps every matching row whole
  {
    auto ts = g.device(0, n);
    Schema ks;
    ks.fields = {Field{"v", FPDB_TYPE_INT64, false}};
    auto keys = TupleSet::fromHost(defaultContext(), ks, {hostInt64({1, 2, 3, 500, 999})});
    auto bf = ArrowBloomFilter::make(5, {"v"});
    CHECK(bf->build(keys).has_value());
    auto kept = BloomFilterUseKernel::filter(ts, bf, {"v"});
    CHECK(kept.has_value());
    if (kept) {
      auto rows = cellsOf(*kept);
      auto all = cellsOf(ts);
      std::set<std::vector<std::string>> allSet(all.begin(), all.end());
      int64_t matches = 0;
      for (auto& row : all) {
        int64_t v = std::stoll(row[2]);
        matches += v == 1 || v == 2 || v == 3 || v == 500 || v == 999;
      }
      int64_t keptMatches = 0;
      for (auto& row : rows) {
        CHECK(allSet.count(row) == 1);
        int64_t v = std::stoll(row[2]);
        keptMatches += v == 1 || v == 2 || v == 3 || v == 500 || v == 999;
      }
      CHECK(keptMatches == matches);
    }
  }
}

static void testBloomJoin(bool arrowImpl) {
  USE_ARROW_BLOOM_FILTER_IMPL = arrowImpl;
  auto out = runJoin(INNER, true);
  USE_ARROW_BLOOM_FILTER_IMPL = true;
  if (!out) return;
  CHECK(out->numRows() == 4);   // no false negatives: the join result is unchanged
}

static void testGroup() {
  // GroupTest.cpp:137-150: sum(ab) group by aa -> 10: 27, 12: 15
  Executor ex;
  auto a = std::make_shared<SourcePOp>("scan", std::vector<std::shared_ptr<TupleSet>>{table3("a", kA)});
  std::vector<std::shared_ptr<AggregateFunction>> fns = {std::make_shared<Sum>("sum", "ab"),
                                                        std::make_shared<Count>("cnt"),
                                                        std::make_shared<Avg>("avg", "ac")};
  auto group = std::make_shared<GroupPOp>("group", std::vector<std::string>{"aa", "sum", "cnt", "avg"}, 0,
                                          std::vector<std::string>{"aa"}, fns);
  auto collate = std::make_shared<CollatePOp>("collate");
  ex.put(a);
  ex.put(group);
  ex.put(collate);
  connect(a, group);
  connect(group, collate);
  auto r = ex.execute();
  CHECK(r.has_value());
  if (!r) return;
  auto out = collate->tupleSet().value();
  CHECK(out->numRows() == 2 && out->numColumns() == 4);
  auto h = out->toHost();
  CHECK(h[0].int64At(0) == 10 && h[0].int64At(1) == 12);
  CHECK(h[1].int64At(0) == 27 && h[1].int64At(1) == 15);
  CHECK(h[2].int64At(0) == 2 && h[2].int64At(1) == 1);
  CHECK(h[3].doubleAt(0) == 16.0 && h[3].doubleAt(1) == 18.0);   // (16 + 17) / 2 truncates for ints
}

// Arrow 6 row hash of one non-null int64 key: bswap32 of the high 32 bits of
// key * 0xC2B2AE3D27D4EB4F (the 8-byte branch of Hashing::HashFixed).
static uint32_t keyHash(int64_t k) {
  uint32_t hi = (uint32_t)(((uint64_t)k * 0xC2B2AE3D27D4EB4FULL) >> 32);
  return (hi >> 24) | ((hi >> 8) & 0xff00u) | ((hi << 8) & 0xff0000u) | (hi << 24);
}

static void testShuffle() {
  // ShufflePOp with three consumers: slot p gets the rows with hash % 3 == p,
  // in input order; an empty input reaches every slot
  const int64_t n = 5000;
  std::vector<int64_t> k(n), v(n);
  for (int64_t i = 0; i < n; ++i) {
    k[i] = (i * 7919) % 1013 - 300;
    v[i] = i;
  }
  Schema s;
  s.fields = {Field{"k", FPDB_TYPE_INT64, false}, Field{"v", FPDB_TYPE_INT64, false}};
  auto ts = TupleSet::fromHost(defaultContext(), s, {hostInt64(k), hostInt64(v)});
  Executor ex;
  auto src = std::make_shared<SourcePOp>("scan", std::vector<std::shared_ptr<TupleSet>>{ts, TupleSet::make(s)});
  auto sh = std::make_shared<ShufflePOp>("shuffle", std::vector<std::string>{"k", "v"}, 0,
                                         std::vector<std::string>{"k"});
  ex.put(src);
  ex.put(sh);
  connect(src, sh);
  std::vector<std::shared_ptr<CollatePOp>> outs;
  for (int p = 0; p < 3; ++p) {
    outs.push_back(std::make_shared<CollatePOp>("collate" + std::to_string(p)));
    ex.put(outs.back());
    connect(sh, outs.back());
  }
  CHECK(sh->getConsumerVec().size() == 3);
  auto r = ex.execute();
  CHECK(r.has_value());
  if (!r) return;
  int64_t total = 0;
  for (int p = 0; p < 3; ++p) {
    auto got = outs[p]->tupleSet();
    CHECK(got.has_value());
    if (!got) continue;
    auto rows = rowsOf(*got);
    std::vector<std::vector<int64_t>> want;
    for (int64_t i = 0; i < n; ++i)
      if (keyHash(k[i]) % 3 == (uint32_t)p) want.push_back({k[i], v[i]});
    CHECK(rows == want);
    total += (int64_t)rows.size();
  }
  CHECK(total == n);
  // ShuffleKernel2 on its own: a missing key column is an error
  auto bad = ShuffleKernel2::shuffle({"nope"}, 2, ts);
  CHECK(!bad.has_value());
}

static void testShuffledJoin() {
  // a two-instance hash join: both sides shuffled on the join key into two
  // slots, one join per slot (the parallel join plans of
  // PrePToPTransformer.cpp), with the same result as one join
  Executor ex;
  auto a = std::make_shared<SourcePOp>("scan_a", std::vector<std::shared_ptr<TupleSet>>{table3("a", kA)});
  auto b = std::make_shared<SourcePOp>("scan_b", std::vector<std::shared_ptr<TupleSet>>{table3("b", kB)});
  auto sa = std::make_shared<ShufflePOp>("shuffle_a", std::vector<std::string>{}, 0, std::vector<std::string>{"aa"});
  auto sb = std::make_shared<ShufflePOp>("shuffle_b", std::vector<std::string>{}, 0, std::vector<std::string>{"ba"});
  auto collate = std::make_shared<CollatePOp>("collate");
  for (auto op : std::vector<std::shared_ptr<PhysicalOp>>{a, b, sa, sb, collate}) ex.put(op);
  connect(a, sa);
  connect(b, sb);
  for (int i = 0; i < 2; ++i) {
    auto join = std::make_shared<HashJoinArrowPOp>("join" + std::to_string(i),
                                                   std::vector<std::string>{"aa", "ab", "ac", "ba", "bb", "bc"}, 0,
                                                   HashJoinPredicate{{"ba"}, {"aa"}}, INNER);
    ex.put(join);
    sb->produce(join);
    join->addBuildProducer(sb);
    sa->produce(join);
    join->addProbeProducer(sa);
    connect(join, collate);
  }
  auto r = ex.execute();
  CHECK(r.has_value());
  if (!r) return;
  auto out = collate->tupleSet().value();
  auto rows = rowsOf(out);
  std::sort(rows.begin(), rows.end());
  std::vector<std::vector<int64_t>> want = {{10, 13, 16, 10, 23, 26}, {10, 13, 16, 10, 25, 28},
                                            {10, 14, 17, 10, 23, 26}, {10, 14, 17, 10, 25, 28}};
  CHECK(rows == want);
}

// One query over partition "lineorder/part0" through the cache: CacheLoadPOp
// -> hit columns to a MergePOp, missed columns scanned from host memory (and
// stored in the cache) -> MergePOp -> collate.
struct CachedScanResult {
  std::shared_ptr<TupleSet> out;
  uint64_t bytesScanned = 0;
};

static CachedScanResult runCachedScan(const std::shared_ptr<SegmentCache>& cache, const Schema& schema,
                                      const std::vector<HostColumn>& cols) {
  Executor ex;
  std::vector<std::string> names;
  for (auto& f : schema.fields) names.push_back(f.name);
  auto actor = std::make_shared<SegmentCacheActor>(cache);
  auto load = std::make_shared<CacheLoadPOp>("cache_load", names, 0, std::vector<std::string>{}, names,
                                             "lineorder/part0", 0, 4096);
  auto scan = std::make_shared<HostScanPOp>("scan", names, 0, schema, cols, "lineorder/part0", 0, 4096,
                                            /*scanOnStart*/ false, /*toCache*/ true);
  auto merge = std::make_shared<MergePOp>("merge", names, 0);
  auto collate = std::make_shared<CollatePOp>("collate");
  for (auto op : std::vector<std::shared_ptr<PhysicalOp>>{actor, load, scan, merge, collate}) ex.put(op);
  load->setHitOperator(merge);
  load->setMissOperatorToCache(scan);
  merge->setLeftProducer(load);
  scan->produce(merge);
  merge->setRightProducer(scan);
  connect(merge, collate);
  auto r = ex.execute();
  CachedScanResult res;
  if (!r) {
    std::printf("  execute failed: %s\n", r.error().c_str());
    ++g_failures;
    return res;
  }
  res.out = collate->tupleSet().value();
  res.bytesScanned = scan->bytesRead();
  return res;
}

static void testSegmentCache() {
  const int64_t n = 3000;
  std::vector<int64_t> a(n), b(n);
  std::vector<int32_t> c(n);
  for (int64_t i = 0; i < n; ++i) {
    a[i] = i * 3 - 7;
    b[i] = (i * 7919) % 1000;
    c[i] = (int32_t)(i % 97);
  }
  Schema s;
  s.fields = {Field{"a", FPDB_TYPE_INT64, false}, Field{"b", FPDB_TYPE_INT64, false},
              Field{"c", FPDB_TYPE_INT32, true}};
  HostColumn hc = hostInt32(c);
  hc.valid.resize(n);
  for (int64_t i = 0; i < n; ++i) hc.valid[i] = i % 5 != 0;
  std::vector<HostColumn> cols = {hostInt64(a), hostInt64(b), hc};
  const uint64_t sizeA = n * 8, sizeC = n * 4 + (n + 7) / 8;
  auto want = rowsOf(TupleSet::fromHost(defaultContext(), s, cols));

  // a cache that holds everything: cold query scans all three columns, warm
  // query scans none; both return the same rows
  auto cache = SegmentCache::make(1 << 20);
  auto cold = runCachedScan(cache, s, cols);
  CHECK(cold.out && rowsOf(cold.out) == want);
  CHECK(cold.bytesScanned == 2 * sizeA + sizeC);
  CHECK(cache->getSize() == 3 && cache->usedBytes() == 2 * sizeA + sizeC);
  CHECK(cache->hitNum() == 0 && cache->missNum() == 3 && cache->shardMissNum() == 1);
  cache->newQuery();
  auto warm = runCachedScan(cache, s, cols);
  CHECK(warm.out && rowsOf(warm.out) == want);
  CHECK(warm.bytesScanned == 0);
  CHECK(cache->hitNum() == 3 && cache->shardHitNum() == 1 && cache->crtQueryHitNum() == 3);
  CHECK(cache->crtQueryMissNum() == 0);

  // room for two of the three segments: storing c evicts a (least recent);
  // the next query hits b and c, scans a again, and the merged columns come
  // back in projection order
  auto small = SegmentCache::make(sizeA + sizeA + sizeC - 1);
  auto q1 = runCachedScan(small, s, cols);
  CHECK(q1.out && rowsOf(q1.out) == want);
  CHECK(small->showCurrentLayout() == "lineorder/part0\tc\nlineorder/part0\tb\n");
  auto q2 = runCachedScan(small, s, cols);
  CHECK(q2.out && rowsOf(q2.out) == want);
  CHECK(q2.bytesScanned == sizeA);
  CHECK(q2.out && q2.out->schema().fields[0].name == "a" && q2.out->schema().fields[2].name == "c");
  CHECK(small->hitNum() == 2 && small->missNum() == 4 && small->shardMissNum() == 2);
}

static void testWireForms() {
  // a blocked filter built from 5000 keys, shipped as its two record batches
  // and restored into a filter of another capacity: same blocks
  std::vector<int64_t> keys(5000);
  for (int64_t i = 0; i < 5000; ++i) keys[i] = i * 7919 % 100003;
  Schema s;
  s.fields = {Field{"k", FPDB_TYPE_INT64, false}};
  auto ts = TupleSet::fromHost(defaultContext(), s, {hostInt64(keys)});
  auto f = ArrowBloomFilter::make(5000, {"k"});
  CHECK(f->build(ts).has_value());
  auto batches = f->makeBitmapRecordBatches();
  CHECK(batches.has_value());
  if (!batches) return;
  CHECK(batches->size() == 2 && (*batches)[0].bytes.size() == 136 &&
        (*batches)[1].bytes.size() == ((size_t)8 << f->logNumBlocks()));
  auto g = ArrowBloomFilter::make(10, {"k"});
  CHECK(g->saveBitmapRecordBatches(*batches).has_value());
  CHECK(g->logNumBlocks() == f->logNumBlocks() && g->blocksToHost() == f->blocksToHost());
  auto bad = *batches;
  bad[0].bytes[0] ^= 1;
  CHECK(!g->saveBitmapRecordBatches(bad).has_value());
  CHECK(!g->saveBitmapRecordBatches({(*batches)[1]}).has_value());

  // the legacy filter's words as one int64 batch
  auto lf = BloomFilter::make(5000, 0.01);
  lf->init(11);
  CHECK(lf->build(ts, {"k"}).has_value());
  auto lb = lf->makeBitmapRecordBatches();
  CHECK(lb.has_value() && lb->size() == 1 && (*lb)[0].type == FPDB_TYPE_INT64);
  auto lg = BloomFilter::make(5000, 0.01);
  lg->init(11);
  CHECK(lg->saveBitmapRecordBatches(*lb).has_value() && lg->getBitArray() == lf->getBitArray());

  // fromJson of each filter's toJson, then its bits from the record batches:
  // the same description, the same bits (BloomFilterBase::fromJson,
  // BloomFilterBase.cpp:25-43; the bits travel separately, BloomFilter.cpp:158)
  {
    auto lj = BloomFilterBase::fromJson(lf->toJson());
    CHECK(lj.has_value());
    if (lj) {
      CHECK((*lj)->getType() == BLOOM_FILTER && (*lj)->toJson() == lf->toJson());
      auto lr = std::static_pointer_cast<BloomFilter>(*lj);
      CHECK(lr->numBits() == lf->numBits() && lr->numHashFunctions() == lf->numHashFunctions() &&
            lr->falsePositiveRate() == 0.01);
      CHECK(lr->saveBitmapRecordBatches(*lb).has_value() && lr->getBitArray() == lf->getBitArray());
    }
    auto aj = BloomFilterBase::fromJson(f->toJson());
    CHECK(aj.has_value());
    if (aj) {
      CHECK((*aj)->getType() == ARROW_BLOOM_FILTER && (*aj)->toJson() == f->toJson());
      auto ar = std::static_pointer_cast<ArrowBloomFilter>(*aj);
      CHECK(ar->columnNames() == std::vector<std::string>{"k"} && ar->logNumBlocks() == f->logNumBlocks());
      CHECK(ar->saveBitmapRecordBatches(*batches).has_value() && ar->blocksToHost() == f->blocksToHost());
    }
    // the reference's field names, written by hand
    auto hand = BloomFilterBase::fromJson(
        "{\"type\": 0, \"capacity\": 10, \"falsePositiveRate\": 0.5, \"valid\": true, \"numHashFunctions\": 1,"
        " \"numBits\": 15, \"hashFunctions\": [{\"a\": 3, \"b\": 4, \"m\": 15, \"p\": 17}]}");
    CHECK(hand.has_value() && std::static_pointer_cast<BloomFilter>(*hand)->hashFunctions().size() == 1);
    CHECK(BloomFilterBase::fromJson("{\"type\": 1, \"capacity\": 10, \"columnNames\": [\"x\"],"
                                    " \"blockedBloomFilter\": {\"log_num_blocks\": 3}}")
              .has_value());
    CHECK(!BloomFilterBase::fromJson("{\"capacity\": 10}").has_value());                  // no type
    CHECK(!BloomFilterBase::fromJson("{\"type\": 7}").has_value());                       // unknown type
    CHECK(!BloomFilterBase::fromJson("{\"type\": 0, \"capacity\": 10}").has_value());     // fields missing
    CHECK(!BloomFilterBase::fromJson("{\"type\": 0, \"capacity\": 10, \"falsePositiveRate\": 0.5, \"valid\": true,"
                                     " \"numHashFunctions\": 1, \"numBits\": 15,"
                                     " \"hashFunctions\": [{\"a\": 3, \"b\": 4, \"m\": 99, \"p\": 101}]}")
               .has_value());                                                             // m beyond the bits
    CHECK(!BloomFilterBase::fromJson("{\"type\": 1, \"capacity\": 10").has_value());      // truncated text
  }

  // a filter's selection over three batches as bitmap words, then per batch back
  fpdb_context* ctx = defaultContext();
  std::vector<int64_t> sizes = {70, 0, 200};
  std::vector<std::vector<uint8_t>> sel;
  FilterBitmap fb;
  int64_t row = 0;
  std::vector<int64_t> want(5, 0);
  for (int64_t n : sizes) {
    std::vector<uint8_t> bits((n + 63) / 64 * 8 + 8, 0);
    for (int64_t i = 0; i < n; ++i)
      if ((i * 37 + n) % 5 == 0) {
        bits[i / 8] |= (uint8_t)(1u << (i % 8));
        want[(row + i) / 64] |= (int64_t)(1ULL << ((row + i) % 64));
      }
    DeviceBuffer d(ctx, bits.size());
    checkStatus(fpdb_copy_to_device(ctx, d.data(), bits.data(), bits.size()), "copy");
    CHECK(fb.bufferBitMap(static_cast<uint8_t*>(d.data()), row, n).has_value());
    sel.push_back(bits);
    row += n;
  }
  want.resize((row + 63) / 64);
  CHECK(fb.toHost() == want);
  auto rb = fb.toRecordBatch();
  auto back = FilterBitmap::fromRecordBatch(rb);
  CHECK(back.has_value());
  if (!back) return;
  row = 0;
  for (size_t b = 0; b < sizes.size(); ++b) {
    const int64_t n = sizes[b];
    auto slice = back->makeSelectionBitmap(row, n);
    CHECK(slice.has_value());
    if (slice && n) {
      std::vector<uint8_t> got((n + 7) / 8);
      checkStatus(fpdb_copy_to_host(ctx, got.data(), (*slice)->data(), got.size()), "copy");
      CHECK(std::equal(got.begin(), got.end(), sel[b].begin()));
    }
    row += n;
  }
}

static void testTwoPhaseGroupBy() {
  // the parallel group-by plan (PrePToPTransformer.cpp:434-560): a partial
  // GroupPOp per input partition, a ShufflePOp on the group key, a reducing
  // GroupPOp per slot (Sum of sums, Count becomes a Sum of counts, Avg becomes
  // a partial Sum + Count and an AvgReduce, PrePToPTransformerUtil.cpp:79-118);
  // the same groups as one GroupPOp over all rows
  const int64_t n = 20000;
  const int parts = 4, slots = 3;
  std::vector<int64_t> k(n), v(n);
  for (int64_t i = 0; i < n; ++i) {
    k[i] = (i * 7919) % 257;
    v[i] = (i * 31) % 1000 - 500;
  }
  Schema s;
  s.fields = {Field{"k", FPDB_TYPE_INT64, false}, Field{"v", FPDB_TYPE_INT64, false}};
  std::map<int64_t, std::pair<int64_t, int64_t>> want;
  for (int64_t i = 0; i < n; ++i) {
    want[k[i]].first += v[i];
    want[k[i]].second += 1;
  }
  Executor ex;
  auto collate = std::make_shared<CollatePOp>("collate");
  ex.put(collate);
  std::vector<std::shared_ptr<GroupPOp>> reducers;
  for (int r = 0; r < slots; ++r) {
    std::vector<std::shared_ptr<AggregateFunction>> fns = {std::make_shared<Sum>("sum_v", "sum_v"),
                                                          std::make_shared<Sum>("cnt", "cnt"),
                                                          std::make_shared<AvgReduce>("avg_v")};
    reducers.push_back(std::make_shared<GroupPOp>("reduce" + std::to_string(r),
                                                  std::vector<std::string>{"k", "sum_v", "cnt", "avg_v"}, 0,
                                                  std::vector<std::string>{"k"}, fns));
    ex.put(reducers.back());
    connect(reducers.back(), collate);
  }
  for (int p = 0; p < parts; ++p) {
    const int64_t b = n * p / parts, e = n * (p + 1) / parts;
    auto ts = TupleSet::fromHost(defaultContext(), s,
                                 {hostInt64({k.begin() + b, k.begin() + e}), hostInt64({v.begin() + b, v.begin() + e})});
    auto src = std::make_shared<SourcePOp>("scan" + std::to_string(p), std::vector<std::shared_ptr<TupleSet>>{ts});
    std::vector<std::shared_ptr<AggregateFunction>> fns = {std::make_shared<Sum>("sum_v", "v"),
                                                          std::make_shared<Count>("cnt")};
    for (auto& f : makeAvgPartial("avg_v", "v")) fns.push_back(f);
    auto partial = std::make_shared<GroupPOp>(
        "partial" + std::to_string(p),
        std::vector<std::string>{"k", "sum_v", "cnt", "AVG_INTERMEDIATE_SUM_avg_v", "AVG_INTERMEDIATE_COUNT_avg_v"}, 0,
        std::vector<std::string>{"k"}, fns);
    auto sh = std::make_shared<ShufflePOp>("shuffle" + std::to_string(p), std::vector<std::string>{}, 0,
                                           std::vector<std::string>{"k"});
    for (auto op : std::vector<std::shared_ptr<PhysicalOp>>{src, partial, sh}) ex.put(op);
    connect(src, partial);
    connect(partial, sh);
    for (auto& r : reducers) connect(sh, r);
  }
  auto r = ex.execute();
  CHECK(r.has_value());
  if (!r) return;
  auto out = collate->tupleSet().value();
  auto h = out->toHost();
  std::map<int64_t, std::pair<int64_t, int64_t>> got;
  CHECK(out->schema().fields.size() == 4 && out->schema().fields[3].name == "avg_v" &&
        out->schema().fields[3].type == FPDB_TYPE_FLOAT64);
  for (int64_t i = 0; i < out->numRows(); ++i) {
    const int64_t key = h[0].int64At(i);
    CHECK(got.count(key) == 0);     // each key reduced in exactly one slot
    got[key] = {h[1].int64At(i), h[2].int64At(i)};
    // AVG of int64 values: integer quotient of the sums, then float64
    const auto& w = want[key];
    CHECK(h[3].isValid(i) && h[3].doubleAt(i) == (double)(w.first / w.second));
  }
  CHECK(got == want);
}

static void testEmptyAndErrors() {
  // an empty build side still yields an (empty) result with the output schema
  Executor ex;
  Schema sb;
  sb.fields = {Field{"ba", FPDB_TYPE_INT64, false}};
  auto emptyB = TupleSet::make(sb);
  auto a = std::make_shared<SourcePOp>("scan_a", std::vector<std::shared_ptr<TupleSet>>{table3("a", kA)});
  auto b = std::make_shared<SourcePOp>("scan_b", std::vector<std::shared_ptr<TupleSet>>{emptyB});
  auto join = std::make_shared<HashJoinArrowPOp>("join", std::vector<std::string>{"aa", "ba"}, 0,
                                                 HashJoinPredicate{{"ba"}, {"aa"}}, INNER);
  auto collate = std::make_shared<CollatePOp>("collate");
  ex.put(a);
  ex.put(b);
  ex.put(join);
  ex.put(collate);
  b->produce(join);
  join->addBuildProducer(b);
  a->produce(join);
  join->addProbeProducer(a);
  connect(join, collate);
  auto r = ex.execute();
  CHECK(r.has_value());
  if (r) {
    auto out = collate->tupleSet().value();
    CHECK(out->numRows() == 0 && out->numColumns() == 2);
  }
  // a bad join column name is reported as an operator error, not a crash
  Executor ex2;
  auto a2 = std::make_shared<SourcePOp>("scan_a", std::vector<std::shared_ptr<TupleSet>>{table3("a", kA)});
  auto b2 = std::make_shared<SourcePOp>("scan_b", std::vector<std::shared_ptr<TupleSet>>{table3("b", kB)});
  auto j2 = std::make_shared<HashJoinArrowPOp>("join", std::vector<std::string>{"aa"}, 0,
                                               HashJoinPredicate{{"nope"}, {"aa"}}, INNER);
  auto c2 = std::make_shared<CollatePOp>("collate");
  ex2.put(a2);
  ex2.put(b2);
  ex2.put(j2);
  ex2.put(c2);
  b2->produce(j2);
  j2->addBuildProducer(b2);
  a2->produce(j2);
  j2->addProbeProducer(a2);
  connect(j2, c2);
  auto r2 = ex2.execute();
  CHECK(!r2.has_value());
  if (!r2) CHECK(r2.error().find("nope") != std::string::npos);
}

// ---- star join ---------------------------------------------------------------

// SplitMix64 over seed ^ column*phi ^ row, as flexpushdowndb_amd/datagen.py
static uint64_t splitmix(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static int64_t uni(uint64_t seed, uint64_t col, int64_t row, int64_t lo, int64_t hi) {
  uint64_t u = splitmix(seed ^ (col * 0x9E3779B97F4A7C15ULL) ^ (uint64_t)row);
  return lo + (int64_t)(u % (uint64_t)(hi - lo + 1));
}

// fact(fk1, fk2, m) x dim1(k1, a1, g1) x dim2(k2, a2, g2):
//   dims filtered by a1 < 30, a2 < 10; forward bloom filters dim -> fact,
//   backward filter fact.fk1 -> dim1; joins; group by (g1, g2): sum(m), count(*).
static int starJoin(int64_t n, const char* outPath) {
  const int64_t d1 = 1000, d2 = 500;
  const uint64_t seed = 77;
  fpdb_context* ctx = defaultContext();
  std::vector<int64_t> fk1(n), fk2(n), m(n), k1(d1), k2(d2);
  std::vector<int32_t> a1(d1), g1(d1), a2(d2), g2(d2);
  for (int64_t i = 0; i < n; ++i) {
    fk1[i] = uni(seed, 1, i, 0, d1 - 1);
    fk2[i] = uni(seed, 2, i, 0, d2 - 1);
    m[i] = uni(seed, 3, i, 1, 100);
  }
  for (int64_t i = 0; i < d1; ++i) {
    k1[i] = i;
    a1[i] = (int32_t)uni(seed, 4, i, 0, 99);
    g1[i] = (int32_t)uni(seed, 5, i, 0, 9);
  }
  for (int64_t i = 0; i < d2; ++i) {
    k2[i] = i;
    a2[i] = (int32_t)uni(seed, 6, i, 0, 99);
    g2[i] = (int32_t)uni(seed, 7, i, 0, 4);
  }
  Schema sf{{Field{"fk1", FPDB_TYPE_INT64}, Field{"fk2", FPDB_TYPE_INT64}, Field{"m", FPDB_TYPE_INT64}}};
  Schema s1{{Field{"k1", FPDB_TYPE_INT64}, Field{"a1", FPDB_TYPE_INT32}, Field{"g1", FPDB_TYPE_INT32}}};
  Schema s2{{Field{"k2", FPDB_TYPE_INT64}, Field{"a2", FPDB_TYPE_INT32}, Field{"g2", FPDB_TYPE_INT32}}};
  // fact arrives in batches of DefaultChunkSize-like size
  std::vector<std::shared_ptr<TupleSet>> factBatches;
  const int64_t chunk = 100000;
  for (int64_t s = 0; s < n; s += chunk) {
    int64_t e = std::min(n, s + chunk);
    factBatches.push_back(TupleSet::fromHost(
        ctx, sf,
        {hostInt64({fk1.begin() + s, fk1.begin() + e}), hostInt64({fk2.begin() + s, fk2.begin() + e}),
         hostInt64({m.begin() + s, m.begin() + e})}));
  }
  auto dim1 = TupleSet::fromHost(ctx, s1, {hostInt64(k1), hostInt32(a1), hostInt32(g1)});
  auto dim2 = TupleSet::fromHost(ctx, s2, {hostInt64(k2), hostInt32(a2), hostInt32(g2)});

  Executor ex;
  auto scanF = std::make_shared<SourcePOp>("scan_fact", factBatches);
  auto scan1 = std::make_shared<SourcePOp>("scan_dim1", std::vector<std::shared_ptr<TupleSet>>{dim1});
  auto scan2 = std::make_shared<SourcePOp>("scan_dim2", std::vector<std::shared_ptr<TupleSet>>{dim2});
  auto filt1 = std::make_shared<FilterPOp>("filter_dim1", std::vector<std::string>{}, 0, "a1", FPDB_CMP_LT, 30);
  auto filt2 = std::make_shared<FilterPOp>("filter_dim2", std::vector<std::string>{}, 0, "a2", FPDB_CMP_LT, 10);
  // forward: dim1 -> fact.fk1, dim2 -> fact.fk2 (chained)
  auto bfc1 = std::make_shared<BloomFilterCreatePOp>("bfc_dim1", std::vector<std::string>{}, 0, std::vector<std::string>{"k1"});
  auto bfu1 = std::make_shared<BloomFilterUsePOp>("bfu_fact_fk1", std::vector<std::string>{}, 0, std::vector<std::string>{"fk1"});
  auto bfc2 = std::make_shared<BloomFilterCreatePOp>("bfc_dim2", std::vector<std::string>{}, 0, std::vector<std::string>{"k2"});
  auto bfu2 = std::make_shared<BloomFilterUsePOp>("bfu_fact_fk2", std::vector<std::string>{}, 0, std::vector<std::string>{"fk2"});
  // backward: filtered fact.fk1 -> dim1
  auto bfcB = std::make_shared<BloomFilterCreatePOp>("bfc_fact_fk1", std::vector<std::string>{}, 0, std::vector<std::string>{"fk1"});
  auto bfuB = std::make_shared<BloomFilterUsePOp>("bfu_dim1", std::vector<std::string>{}, 0, std::vector<std::string>{"k1"});
  auto join1 = std::make_shared<HashJoinArrowPOp>("join_dim1", std::vector<std::string>{"fk2", "m", "g1"}, 0,
                                                  HashJoinPredicate{{"k1"}, {"fk1"}}, INNER);
  auto join2 = std::make_shared<HashJoinArrowPOp>("join_dim2", std::vector<std::string>{"m", "g1", "g2"}, 0,
                                                  HashJoinPredicate{{"k2"}, {"fk2"}}, INNER);
  std::vector<std::shared_ptr<AggregateFunction>> fns = {std::make_shared<Sum>("sum_m", "m"),
                                                        std::make_shared<Count>("cnt")};
  auto group = std::make_shared<GroupPOp>("group", std::vector<std::string>{"g1", "g2", "sum_m", "cnt"}, 0,
                                          std::vector<std::string>{"g1", "g2"}, fns);
  auto collate = std::make_shared<CollatePOp>("collate");
  for (auto op : std::vector<std::shared_ptr<PhysicalOp>>{scanF, scan1, scan2, filt1, filt2, bfc1, bfu1, bfc2, bfu2,
                                                          bfcB, bfuB, join1, join2, group, collate})
    ex.put(op);
  connect(scan1, filt1);
  connect(scan2, filt2);
  connect(filt1, bfc1);
  bfc1->addBloomFilterUsePOp(bfu1);
  bfu1->consume(bfc1);
  connect(scanF, bfu1);
  connect(filt2, bfc2);
  bfc2->addBloomFilterUsePOp(bfu2);
  bfu2->consume(bfc2);
  connect(bfu1, bfu2);
  connect(bfu2, bfcB);           // fact passes through the backward create
  bfcB->addBloomFilterUsePOp(bfuB);
  bfuB->consume(bfcB);
  connect(bfc1, bfuB);           // dim1 passes through its forward create into the backward use
  bfuB->produce(join1);
  join1->addBuildProducer(bfuB);
  bfcB->produce(join1);
  join1->addProbeProducer(bfcB);
  bfc2->produce(join2);
  join2->addBuildProducer(bfc2);
  join1->produce(join2);
  join2->addProbeProducer(join1);
  connect(join2, group);
  connect(group, collate);
  auto r = ex.execute();
  if (!r) {
    std::printf("starjoin failed: %s\n", r.error().c_str());
    return 1;
  }
  auto out = collate->tupleSet().value();
  FILE* f = std::fopen(outPath, "w");
  if (!f) return 1;
  std::fprintf(f, "g1,g2,sum_m,cnt\n");
  auto h = out->toHost();
  for (int64_t i = 0; i < out->numRows(); ++i)
    std::fprintf(f, "%lld,%lld,%lld,%lld\n", (long long)h[0].int64At(i), (long long)h[1].int64At(i),
                 (long long)h[2].int64At(i), (long long)h[3].int64At(i));
  std::fclose(f);
  auto bfuFact = std::static_pointer_cast<BloomFilterUsePOp>(ex.get("bfu_fact_fk2"));
  std::printf("starjoin rows=%lld groups=%lld fact_after_bloom=%lld\n", (long long)n, (long long)out->numRows(),
              (long long)bfuFact->numRowsOutput());
  return 0;
}

static std::vector<std::string> split(const std::string& s, char sep) {
  std::vector<std::string> out;
  std::string cur;
  for (char ch : s) {
    if (ch == sep) {
      out.push_back(cur);
      cur.clear();
    } else {
      cur += ch;
    }
  }
  if (!cur.empty()) out.push_back(cur);
  return out;
}

// pop_test plan ORDER TABLES JOINS: builds only the predicate-transfer part of
// a plan (no data, no device work) and prints its connection steps.
//   ORDER   s2l | bfs | yan
//   TABLES  name=rows,name=rows,...
//   JOINS   left.c1+c2>right.d1+d2@TYPE;...   (TYPE INNER LEFT RIGHT FULL)
static int planOnly(const std::string& order, const std::string& tablesSpec, const std::string& joinsSpec) {
  Executor ex;
  std::vector<PredTransTable> tables;
  for (const auto& t : split(tablesSpec, ',')) {
    auto kv = split(t, '=');
    auto src = std::make_shared<SourcePOp>("scan_" + kv[0], std::vector<std::shared_ptr<TupleSet>>{});
    ex.put(src);
    tables.push_back(PredTransTable{kv[0], src, std::atof(kv[1].c_str())});
  }
  std::vector<JoinOrigin> joins;
  for (const auto& j : split(joinsSpec, ';')) {
    auto at = split(j, '@');
    auto sides = split(at[0], '>');
    auto l = split(sides[0], '.'), r = split(sides[1], '.');
    JoinOrigin o;
    o.left = l[0];
    o.right = r[0];
    o.leftColumns = split(l[1], '+');
    o.rightColumns = split(r[1], '+');
    const std::string& ty = at[1];
    o.joinType = ty == "LEFT" ? LEFT : ty == "RIGHT" ? RIGHT : ty == "FULL" ? FULL : INNER;
    joins.push_back(o);
  }
  std::vector<PredTransStep> steps;
  auto r = PredTransOrder::orderPredTrans(order == "s2l" ? PredTransOrderType::SMALL_TO_LARGE : PredTransOrderType::BFS,
                                          order == "yan", ex, tables, joins, &steps);
  if (!r) {
    std::printf("ERROR %s\n", r.error().c_str());
    return 0;
  }
  for (const auto& st : steps)
    std::printf("STEP %s|%s|%s|%s|%s|%s\n", st.transferOp.c_str(), st.createOp.c_str(), st.srcTable.c_str(),
                st.tgtTable.c_str(), st.srcAttachedTo.c_str(), st.tgtAttachedTo.c_str());
  for (const auto& kv : *r) std::printf("END %s|%s\n", kv.first.c_str(), kv.second->name().c_str());
  return 0;
}

// The star join of starJoin with its predicate transfer laid out by the plan
// builder (ORDER s2l | bfs | yan) instead of by hand; joins and group-by are
// connected from the chain ends it returns.
static int starJoinPlanned(const std::string& order, int64_t n, const char* outPath) {
  const int64_t d1 = 1000, d2 = 500;
  const uint64_t seed = 77;
  fpdb_context* ctx = defaultContext();
  std::vector<int64_t> fk1(n), fk2(n), m(n), k1(d1), k2(d2);
  std::vector<int32_t> a1(d1), g1(d1), a2(d2), g2(d2);
  for (int64_t i = 0; i < n; ++i) {
    fk1[i] = uni(seed, 1, i, 0, d1 - 1);
    fk2[i] = uni(seed, 2, i, 0, d2 - 1);
    m[i] = uni(seed, 3, i, 1, 100);
  }
  for (int64_t i = 0; i < d1; ++i) {
    k1[i] = i;
    a1[i] = (int32_t)uni(seed, 4, i, 0, 99);
    g1[i] = (int32_t)uni(seed, 5, i, 0, 9);
  }
  for (int64_t i = 0; i < d2; ++i) {
    k2[i] = i;
    a2[i] = (int32_t)uni(seed, 6, i, 0, 99);
    g2[i] = (int32_t)uni(seed, 7, i, 0, 4);
  }
  Schema sf{{Field{"fk1", FPDB_TYPE_INT64}, Field{"fk2", FPDB_TYPE_INT64}, Field{"m", FPDB_TYPE_INT64}}};
  Schema s1{{Field{"k1", FPDB_TYPE_INT64}, Field{"a1", FPDB_TYPE_INT32}, Field{"g1", FPDB_TYPE_INT32}}};
  Schema s2{{Field{"k2", FPDB_TYPE_INT64}, Field{"a2", FPDB_TYPE_INT32}, Field{"g2", FPDB_TYPE_INT32}}};
  std::vector<std::shared_ptr<TupleSet>> factBatches;
  const int64_t chunk = 100000;
  for (int64_t s = 0; s < n; s += chunk) {
    int64_t e = std::min(n, s + chunk);
    factBatches.push_back(TupleSet::fromHost(
        ctx, sf,
        {hostInt64({fk1.begin() + s, fk1.begin() + e}), hostInt64({fk2.begin() + s, fk2.begin() + e}),
         hostInt64({m.begin() + s, m.begin() + e})}));
  }
  auto dim1 = TupleSet::fromHost(ctx, s1, {hostInt64(k1), hostInt32(a1), hostInt32(g1)});
  auto dim2 = TupleSet::fromHost(ctx, s2, {hostInt64(k2), hostInt32(a2), hostInt32(g2)});

  Executor ex;
  auto scanF = std::make_shared<SourcePOp>("scan_fact", factBatches, std::vector<std::string>{"fk1", "fk2", "m"});
  auto scan1 = std::make_shared<SourcePOp>("scan_dim1", std::vector<std::shared_ptr<TupleSet>>{dim1});
  auto scan2 = std::make_shared<SourcePOp>("scan_dim2", std::vector<std::shared_ptr<TupleSet>>{dim2});
  auto filt1 = std::make_shared<FilterPOp>("filter_dim1", std::vector<std::string>{"k1", "a1", "g1"}, 0, "a1",
                                           FPDB_CMP_LT, 30);
  auto filt2 = std::make_shared<FilterPOp>("filter_dim2", std::vector<std::string>{"k2", "a2", "g2"}, 0, "a2",
                                           FPDB_CMP_LT, 10);
  for (auto op : std::vector<std::shared_ptr<PhysicalOp>>{scanF, scan1, scan2, filt1, filt2}) ex.put(op);
  connect(scan1, filt1);
  connect(scan2, filt2);
  std::vector<PredTransTable> tables = {{"fact", scanF, (double)n}, {"dim1", filt1, (double)d1},
                                        {"dim2", filt2, (double)d2}};
  std::vector<JoinOrigin> joins = {{"dim1", "fact", {"k1"}, {"fk1"}, INNER}, {"dim2", "fact", {"k2"}, {"fk2"}, INNER}};
  auto ends = PredTransOrder::orderPredTrans(order == "s2l" ? PredTransOrderType::SMALL_TO_LARGE
                                                            : PredTransOrderType::BFS,
                                             order == "yan", ex, tables, joins);
  if (!ends) {
    std::printf("plan failed: %s\n", ends.error().c_str());
    return 1;
  }
  auto endF = ends->at("fact"), end1 = ends->at("dim1"), end2 = ends->at("dim2");
  auto join1 = std::make_shared<HashJoinArrowPOp>("join_dim1", std::vector<std::string>{"fk2", "m", "g1"}, 0,
                                                  HashJoinPredicate{{"k1"}, {"fk1"}}, INNER);
  auto join2 = std::make_shared<HashJoinArrowPOp>("join_dim2", std::vector<std::string>{"m", "g1", "g2"}, 0,
                                                  HashJoinPredicate{{"k2"}, {"fk2"}}, INNER);
  std::vector<std::shared_ptr<AggregateFunction>> fns = {std::make_shared<Sum>("sum_m", "m"),
                                                        std::make_shared<Count>("cnt")};
  auto group = std::make_shared<GroupPOp>("group", std::vector<std::string>{"g1", "g2", "sum_m", "cnt"}, 0,
                                          std::vector<std::string>{"g1", "g2"}, fns);
  auto collate = std::make_shared<CollatePOp>("collate");
  for (auto op : std::vector<std::shared_ptr<PhysicalOp>>{join1, join2, group, collate}) ex.put(op);
  end1->produce(join1);
  join1->addBuildProducer(end1);
  endF->produce(join1);
  join1->addProbeProducer(endF);
  end2->produce(join2);
  join2->addBuildProducer(end2);
  join1->produce(join2);
  join2->addProbeProducer(join1);
  connect(join2, group);
  connect(group, collate);
  auto t0 = std::chrono::steady_clock::now();
  auto r = ex.execute();
  checkStatus(fpdb_context_synchronize(ctx), "synchronize");
  const double execMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (!r) {
    std::printf("starjoin-planned failed: %s\n", r.error().c_str());
    return 1;
  }
  auto out = collate->tupleSet().value();
  auto h = out->toHost();
  FILE* f = std::fopen(outPath, "w");
  if (!f) return 1;
  std::fprintf(f, "g1,g2,sum_m,cnt\n");
  for (int64_t i = 0; i < out->numRows(); ++i)
    std::fprintf(f, "%lld,%lld,%lld,%lld\n", (long long)h[0].int64At(i), (long long)h[1].int64At(i),
                 (long long)h[2].int64At(i), (long long)h[3].int64At(i));
  std::fclose(f);
  std::printf("starjoin-planned order=%s rows=%lld groups=%lld execute_ms=%.2f\n", order.c_str(), (long long)n,
              (long long)out->numRows(), execMs);
  return 0;
}

int main(int argc, char** argv) {
  std::string mode = argc > 1 ? argv[1] : "unit";
  if (mode == "starjoin-planned") {
    if (argc < 5) return 2;
    return starJoinPlanned(argv[2], std::atoll(argv[3]), argv[4]);
  }
  if (mode == "plan") {
    if (argc < 5) return 2;
    return planOnly(argv[2], argv[3], argv[4]);
  }
  if (mode == "starjoin") {
    if (argc < 4) return 2;
    return starJoin(std::atoll(argv[2]), argv[3]);
  }
  std::vector<std::pair<std::string, std::function<void()>>> tests = {
      {"join", testJoin},
      {"outer_semi_joins", testOuterSemiJoins},
      {"build_probe_join", testBuildProbeJoin},
      {"utf8_and_wide_key_joins", testUtf8AndWideKeyJoins},
      {"utf8_group_filter_shuffle", testUtf8GroupFilterShuffle},
      {"bloom_join_blocked", [] { testBloomJoin(true); }},
      {"bloom_join_legacy", [] { testBloomJoin(false); }},
      {"group", testGroup},
      {"empty_and_errors", testEmptyAndErrors},
      {"shuffle", testShuffle},
      {"shuffled_join", testShuffledJoin},
      {"segment_cache", testSegmentCache},
      {"wire_forms", testWireForms},
      {"two_phase_group_by", testTwoPhaseGroupBy},
  };
  for (auto& t : tests) {
    int before = g_failures;
    std::printf("RUN %s\n", t.first.c_str());
    std::fflush(stdout);
    try {
      t.second();
    } catch (const std::exception& e) {
      std::printf("  exception: %s\n", e.what());
      ++g_failures;
    }
    std::printf("%s %s\n", g_failures == before ? "OK" : "FAIL", t.first.c_str());
    std::fflush(stdout);
  }
  return g_failures ? 1 : 0;
}

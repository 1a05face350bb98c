// Predicate-transfer plan builder: given the base-table chains of a query and
// its base-table joins (JoinOrigin, fpdb-plan/include/fpdb/plan/prephysical/JoinOriginTracer.h:26-50),
// adds the bloom-filter operators that carry predicates across the join graph
// and wires them into each table's chain, as PredTransOrder::orderPredTrans
// (fpdb-executor/src/physical/transform/pred-trans/PredTransOrder.cpp:10-31)
// does with either ordering:
//   SMALL_TO_LARGE  SmallToLargePredTransOrder.cpp:14-214: one forward filter
//                   (left -> right, blocked by RIGHT joins) and one backward
//                   filter (right -> left, blocked by LEFT joins) per join;
//                   each direction connected in a topological order of the
//                   dependency graph (a table's filters are created only after
//                   every filter of that direction has been applied to it);
//                   a cycle is an error.
//   BFS             BFSPredTransOrder.cpp:14-227: a BFS over the join graph
//                   from the largest table (any table for Yannakakis), forward
//                   transfers in reverse BFS order (leaves towards the root),
//                   backward transfers in BFS order; with Yannakakis the
//                   transfers are RIGHT_SEMI hash joins instead of filters.
// FULL joins transfer nothing.  Where the reference iterates hash sets (so the
// order among independent transfers depends on hashing), this builder uses the
// order in which tables and joins are given; results do not depend on it.
#pragma once

#include <map>

#include "BloomFilterPOps.h"
#include "HashJoin.h"

namespace fpdb::amd {

enum class PredTransOrderType { SMALL_TO_LARGE, BFS };

// One base table: the operator its rows leave from (a scan or filter), and its
// estimated row count (used by BFS to pick the root).
struct PredTransTable {
  std::string name;
  std::shared_ptr<PhysicalOp> upConnOp;
  double rowCount = 0;
};

// A join between two base tables; left is the build side (JoinOrigin).
struct JoinOrigin {
  std::string left, right;
  std::vector<std::string> leftColumns, rightColumns;
  JoinType joinType = INNER;
};

// One connected transfer, in connection order: the operator that transfers
// (a bloom-filter use, or a semi-join), the one that reads the source table
// (the filter's create; the semi-join itself), and the operators each side
// was attached to.
struct PredTransStep {
  std::string transferOp, createOp;
  std::string srcTable, tgtTable;
  std::string srcAttachedTo, tgtAttachedTo;
};

class PredTransOrder {
 public:
  // Adds the transfer operators to `ex` and returns, per table, the operator
  // its (filtered) rows now leave from (PredTransOrder::updateTransRes,
  // PredTransOrder.cpp:41-61).  Tables named by no join keep their operator.
  // `steps` (optional) receives the transfers in the order they were connected.
  static Expected<std::map<std::string, std::shared_ptr<PhysicalOp>>> orderPredTrans(
      PredTransOrderType type, bool yannakakis, Executor& ex, const std::vector<PredTransTable>& tables,
      const std::vector<JoinOrigin>& joinOrigins, std::vector<PredTransStep>* steps = nullptr);
};

}  // namespace fpdb::amd

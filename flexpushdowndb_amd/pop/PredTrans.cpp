#include "PredTrans.h"

#include <deque>
#include <stdexcept>

namespace fpdb::amd {

namespace {

// HashJoinPredicate::toString (fpdb-executor/src/physical/join/hashjoin/HashJoinPredicate.cpp:36-47)
std::string predicateString(const std::vector<std::string>& left, const std::vector<std::string>& right) {
  std::string s;
  for (size_t i = 0; i < left.size(); ++i) {
    s += left[i] + "-" + right[i];
    if (i + 1 < left.size()) s += "-";
  }
  return s;
}

struct Unit {
  std::string name;
  std::shared_ptr<PhysicalOp> origUpConnOp, currUpConnOp;
  double rowCount = 0;
  // SMALL_TO_LARGE dependency bookkeeping
  std::vector<int> fwOut, bwOut;
  int numFwToVisit = 0, numBwToVisit = 0, numFwVisited = 0, numBwVisited = 0;
  // BFS neighbours: (unit, transfer direction from this unit to it, this unit's columns, its columns)
  struct Neighbor {
    int unit;
    int dir;   // 0 both, 1 forward only, 2 backward only
    std::vector<std::string> ownColumns, otherColumns;
  };
  std::vector<Neighbor> neighbors;
};

struct Planner {
  Executor& ex;
  std::vector<Unit> units;
  std::map<std::string, int> byName;
  unsigned idGen = 0;
  std::vector<PredTransStep> steps;

  explicit Planner(Executor& e) : ex(e) {}

  Expected<int> unitOf(const std::string& name) {
    auto it = byName.find(name);
    if (it == byName.end()) return make_unexpected("predicate transfer: unknown table '" + name + "'");
    return it->second;
  }

  void link(const std::shared_ptr<PhysicalOp>& producer, const std::shared_ptr<PhysicalOp>& consumer) {
    producer->produce(consumer);
    consumer->consume(producer);
  }

  // A bloom-filter create on `src`'s current operator and its use on `tgt`'s;
  // the use becomes `tgt`'s current operator.
  void connectFilter(const std::shared_ptr<BloomFilterCreatePOp>& create, const std::shared_ptr<BloomFilterUsePOp>& use,
                     int src, int tgt) {
    steps.push_back(PredTransStep{use->name(), create->name(), units[src].name, units[tgt].name,
                                  units[src].currUpConnOp->name(), units[tgt].currUpConnOp->name()});
    link(units[src].currUpConnOp, create);
    link(units[tgt].currUpConnOp, use);
    units[tgt].currUpConnOp = use;
  }

  std::pair<std::shared_ptr<BloomFilterCreatePOp>, std::shared_ptr<BloomFilterUsePOp>> makeFilter(
      const std::string& dirTag, unsigned id, const std::string& pred, const std::vector<std::string>& createColumns,
      const std::vector<std::string>& useColumns) {
    auto create = std::make_shared<BloomFilterCreatePOp>("BloomFilterCreate(" + dirTag + ")<" + std::to_string(id) +
                                                             ">-" + pred,
                                                         std::vector<std::string>{}, 0, createColumns);
    auto use = std::make_shared<BloomFilterUsePOp>("BloomFilterUse(" + dirTag + ")<" + std::to_string(id) + ">-" + pred,
                                                   std::vector<std::string>{}, 0, useColumns);
    create->addBloomFilterUsePOp(use);
    use->consume(create);
    ex.put(create);
    ex.put(use);
    return {create, use};
  }

  // ---- SMALL_TO_LARGE (SmallToLargePredTransOrder.cpp:37-214) ----
  struct Node {
    std::shared_ptr<BloomFilterCreatePOp> create;
    std::shared_ptr<BloomFilterUsePOp> use;
    int createUnit, useUnit;
  };

  Expected<void> smallToLarge(const std::vector<JoinOrigin>& joins) {
    std::vector<Node> fw, bw;
    for (const auto& j : joins) {
      auto l = unitOf(j.left);
      if (!l) return make_unexpected(l.error());
      auto r = unitOf(j.right);
      if (!r) return make_unexpected(r.error());
      if (j.joinType == FULL) continue;
      const std::string pred = predicateString(j.leftColumns, j.rightColumns);
      const unsigned id = idGen++;
      if (j.joinType != RIGHT) {
        auto [c, u] = makeFilter("F", id, pred, j.leftColumns, j.rightColumns);
        units[*l].fwOut.push_back((int)fw.size());
        ++units[*r].numFwToVisit;
        fw.push_back(Node{c, u, *l, *r});
      }
      if (j.joinType != LEFT) {
        auto [c, u] = makeFilter("B", id, pred, j.rightColumns, j.leftColumns);
        units[*r].bwOut.push_back((int)bw.size());
        ++units[*l].numBwToVisit;
        bw.push_back(Node{c, u, *r, *l});
      }
    }
    if (auto s = connectTopological(fw, true); !s) return s;
    return connectTopological(bw, false);
  }

  Expected<void> connectTopological(const std::vector<Node>& nodes, bool forward) {
    std::vector<bool> pending(nodes.size(), true);
    std::deque<int> freeNodes;
    for (size_t i = 0; i < nodes.size(); ++i) {
      const Unit& cu = units[nodes[i].createUnit];
      if ((forward ? cu.numFwToVisit : cu.numBwToVisit) == 0) {
        freeNodes.push_back((int)i);
        pending[i] = false;
      }
    }
    while (!freeNodes.empty()) {
      const Node& node = nodes[freeNodes.front()];
      freeNodes.pop_front();
      connectFilter(node.create, node.use, node.createUnit, node.useUnit);
      Unit& uu = units[node.useUnit];
      int& visited = forward ? uu.numFwVisited : uu.numBwVisited;
      ++visited;
      if (visited == (forward ? uu.numFwToVisit : uu.numBwToVisit)) {
        for (int out : forward ? uu.fwOut : uu.bwOut) {
          freeNodes.push_back(out);
          pending[out] = false;
        }
      }
    }
    for (bool p : pending)
      if (p)
        return make_unexpected(std::string("The join origins (base table joins) contain cycles (") +
                               (forward ? "forward" : "backward") + " predicate transfer)");
    return {};
  }

  // ---- BFS (BFSPredTransOrder.cpp:14-227) ----
  struct Pair {
    int src, tgt;
    std::vector<std::string> srcColumns, tgtColumns;
    bool forward, backward;
  };

  Expected<void> bfs(const std::vector<JoinOrigin>& joins, bool yannakakis) {
    std::vector<bool> inGraph(units.size(), false);
    for (const auto& j : joins) {
      auto l = unitOf(j.left);
      if (!l) return make_unexpected(l.error());
      auto r = unitOf(j.right);
      if (!r) return make_unexpected(r.error());
      if (j.joinType == FULL) continue;
      int leftDir = 0, rightDir = 0;
      if (j.joinType == LEFT) {
        leftDir = 1;
        rightDir = 2;
      } else if (j.joinType == RIGHT) {
        leftDir = 2;
        rightDir = 1;
      }
      units[*l].neighbors.push_back({*r, leftDir, j.leftColumns, j.rightColumns});
      units[*r].neighbors.push_back({*l, rightDir, j.rightColumns, j.leftColumns});
      inGraph[*l] = inGraph[*r] = true;
    }
    // BFS from the largest remaining table (the first remaining one for
    // Yannakakis) until every table in the join graph is visited
    std::vector<bool> toVisit = inGraph;
    std::vector<Pair> forwardStack;
    for (;;) {
      int root = -1;
      for (size_t u = 0; u < units.size(); ++u) {
        if (!toVisit[u]) continue;
        if (root < 0 || (!yannakakis && units[u].rowCount > units[root].rowCount)) root = (int)u;
        if (yannakakis) break;
      }
      if (root < 0) break;
      std::deque<int> q{root};
      toVisit[root] = false;
      while (!q.empty()) {
        const int u = q.front();
        q.pop_front();
        for (const auto& nb : units[u].neighbors) {
          if (!toVisit[nb.unit]) continue;
          const bool forward = nb.dir == 0 || nb.dir == 1;
          const bool backward = nb.dir == 0 || nb.dir == 2;
          // transfer from the newly reached table towards the one it was reached from
          forwardStack.push_back(Pair{nb.unit, u, nb.otherColumns, nb.ownColumns, backward, forward});
          q.push_back(nb.unit);
          toVisit[nb.unit] = false;
        }
      }
    }
    std::vector<Pair> backwardStack;
    while (!forwardStack.empty()) {
      Pair p = forwardStack.back();
      forwardStack.pop_back();
      if (p.forward) transfer(p, "F", yannakakis);
      backwardStack.push_back(Pair{p.tgt, p.src, p.tgtColumns, p.srcColumns, p.backward, p.forward});
    }
    while (!backwardStack.empty()) {
      Pair p = backwardStack.back();
      backwardStack.pop_back();
      if (p.forward) transfer(p, "B", yannakakis);
    }
    return {};
  }

  void transfer(const Pair& p, const std::string& dirTag, bool yannakakis) {
    const std::string pred = predicateString(p.srcColumns, p.tgtColumns);
    const unsigned id = idGen++;
    if (yannakakis) {
      auto join = std::make_shared<HashJoinArrowPOp>(
          "HashJoinArrow(" + dirTag + ")<" + std::to_string(id) + ">-" + pred,
          units[p.tgt].origUpConnOp->getProjectColumnNames(), 0, HashJoinPredicate{p.srcColumns, p.tgtColumns},
          RIGHT_SEMI);
      steps.push_back(PredTransStep{join->name(), join->name(), units[p.src].name, units[p.tgt].name,
                                    units[p.src].currUpConnOp->name(), units[p.tgt].currUpConnOp->name()});
      join->addBuildProducer(units[p.src].currUpConnOp);
      join->addProbeProducer(units[p.tgt].currUpConnOp);
      units[p.src].currUpConnOp->produce(join);
      units[p.tgt].currUpConnOp->produce(join);
      ex.put(join);
      units[p.tgt].currUpConnOp = join;
    } else {
      auto [c, u] = makeFilter(dirTag, id, pred, p.srcColumns, p.tgtColumns);
      connectFilter(c, u, p.src, p.tgt);
    }
  }
};

}  // namespace

Expected<std::map<std::string, std::shared_ptr<PhysicalOp>>> PredTransOrder::orderPredTrans(
    PredTransOrderType type, bool yannakakis, Executor& ex, const std::vector<PredTransTable>& tables,
    const std::vector<JoinOrigin>& joinOrigins, std::vector<PredTransStep>* steps) {
  Planner pl(ex);
  for (const auto& t : tables) {
    if (pl.byName.count(t.name)) return make_unexpected("predicate transfer: duplicate table '" + t.name + "'");
    pl.byName[t.name] = (int)pl.units.size();
    Unit u;
    u.name = t.name;
    u.origUpConnOp = u.currUpConnOp = t.upConnOp;
    u.rowCount = t.rowCount;
    pl.units.push_back(std::move(u));
  }
  Expected<void> r = type == PredTransOrderType::SMALL_TO_LARGE ? pl.smallToLarge(joinOrigins)
                                                                 : pl.bfs(joinOrigins, yannakakis);
  if (!r) return make_unexpected(r.error());
  std::map<std::string, std::shared_ptr<PhysicalOp>> out;
  for (const auto& u : pl.units) out[u.name] = u.currUpConnOp;
  if (steps) *steps = std::move(pl.steps);
  return out;
}

}  // namespace fpdb::amd

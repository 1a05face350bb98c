// Physical operator base, operator context and a single-threaded in-process
// mailbox runtime.  Mirrors PhysicalOp (fpdb-executor/include/fpdb/executor/
// physical/PhysicalOp.h:26-108), POpContext (fpdb-executor/src/physical/
// POpContext.cpp:25-195) and the START/TUPLESET/BLOOM_FILTER/COMPLETE/ERROR
// protocol of POpActor (fpdb-executor/src/physical/POpActor.cpp:38-135),
// delivering messages in FIFO order like CAF with one worker thread.
#pragma once

#include <deque>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "Expected.h"
#include "Message.h"

namespace fpdb::amd {

enum class POpType {
  HASH_JOIN_ARROW, HASH_JOIN_BUILD, HASH_JOIN_PROBE, BLOOM_FILTER_CREATE, BLOOM_FILTER_USE, GROUP, SHUFFLE, SOURCE, COLLATE, FILTER,
  CACHE_LOAD, SEGMENT_CACHE, FILE_SCAN, MERGE
};

enum class POpRelationshipType { Producer, Consumer };

class Executor;
class PhysicalOp;

class OperatorMap {
 public:
  void addProducer(const std::string& name) { producerComplete_[name] = false; }
  void setComplete(const std::string& name) {
    auto it = producerComplete_.find(name);
    if (it != producerComplete_.end()) it->second = true;
  }
  bool allComplete(POpRelationshipType) const {
    for (auto& kv : producerComplete_)
      if (!kv.second) return false;
    return true;
  }

 private:
  std::map<std::string, bool> producerComplete_;
};

class POpContext {
 public:
  POpContext(PhysicalOp* op, Executor* exec) : op_(op), exec_(exec) {}
  // Sends to every consumer, or to the given recipients.
  void tell(const std::shared_ptr<Message>& msg);
  void tell(const std::shared_ptr<Message>& msg, const std::set<std::string>& recipients);
  void notifyComplete();
  void notifyError(const std::string& content);
  bool isComplete() const { return complete_; }
  OperatorMap& operatorMap() { return operatorMap_; }

 private:
  PhysicalOp* op_;
  Executor* exec_;
  bool complete_ = false;
  OperatorMap operatorMap_;
};

class PhysicalOp {
 public:
  PhysicalOp(std::string name, POpType type, std::vector<std::string> projectColumnNames, int nodeId)
      : name_(std::move(name)), type_(type), projectColumnNames_(std::move(projectColumnNames)), nodeId_(nodeId) {}
  virtual ~PhysicalOp() = default;

  std::string& name() { return name_; }
  POpType getType() const { return type_; }
  virtual std::string getTypeString() const = 0;
  const std::vector<std::string>& getProjectColumnNames() const { return projectColumnNames_; }
  int getNodeId() const { return nodeId_; }
  const std::set<std::string>& producers() const { return producers_; }
  const std::set<std::string>& consumers() const { return consumers_; }

  virtual void produce(const std::shared_ptr<PhysicalOp>& op) { consumers_.insert(op->name()); }
  virtual void consume(const std::shared_ptr<PhysicalOp>& op) { producers_.insert(op->name()); }

  void create(const std::shared_ptr<POpContext>& ctx) { opContext_ = ctx; }
  std::shared_ptr<POpContext> ctx() { return opContext_; }

  virtual void onReceive(const Envelope& msg) = 0;
  virtual void clear() = 0;

 protected:
  std::string name_;
  POpType type_;
  std::vector<std::string> projectColumnNames_;
  int nodeId_;
  std::set<std::string> producers_;
  std::set<std::string> consumers_;
  std::shared_ptr<POpContext> opContext_;
};

// Connects producer -> consumer in both directions.
void connect(const std::shared_ptr<PhysicalOp>& producer, const std::shared_ptr<PhysicalOp>& consumer);

// Runs a plan to completion: START to every operator, then messages in FIFO
// order until every operator has completed.  The first ERROR aborts the run
// and is returned (Execution::join, fpdb-executor/src/Execution.cpp:216-290).
class Executor {
 public:
  void put(const std::shared_ptr<PhysicalOp>& op);
  std::shared_ptr<PhysicalOp> get(const std::string& name) const;
  Expected<void> execute();
  void enqueue(const std::string& recipient, const std::shared_ptr<Message>& msg);
  void reportError(const std::string& sender, const std::string& content);
  // Wall time spent inside each operator's onReceive (POpActor.cpp:40-97).
  const std::map<std::string, double>& processingTimeMs() const { return processingMs_; }

 private:
  std::map<std::string, std::shared_ptr<PhysicalOp>> ops_;
  std::vector<std::string> order_;
  std::deque<std::pair<std::string, std::shared_ptr<Message>>> queue_;
  std::string error_;
  bool failed_ = false;
  std::map<std::string, double> processingMs_;
};

}  // namespace fpdb::amd

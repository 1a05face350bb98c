#include "PhysicalOp.h"

#include <chrono>
#include <stdexcept>

namespace fpdb::amd {

std::string Message::getTypeString() const {
  switch (type_) {
    case MessageType::START: return "StartMessage";
    case MessageType::TUPLESET: return "TupleSetMessage";
    case MessageType::TUPLESET_INDEX: return "TupleSetIndexMessage";
    case MessageType::BLOOM_FILTER: return "BloomFilterMessage";
    case MessageType::COMPLETE: return "CompleteMessage";
    case MessageType::ERROR: return "ErrorMessage";
    case MessageType::SCAN: return "ScanMessage";
    case MessageType::LOAD_REQUEST: return "LoadRequestMessage";
    case MessageType::LOAD_RESPONSE: return "LoadResponseMessage";
    case MessageType::STORE_REQUEST: return "StoreRequestMessage";
    case MessageType::CACHE_METRICS: return "CacheMetricsMessage";
  }
  return "Message";
}

void POpContext::tell(const std::shared_ptr<Message>& msg) { tell(msg, op_->consumers()); }

void POpContext::tell(const std::shared_ptr<Message>& msg, const std::set<std::string>& recipients) {
  for (auto& r : recipients) exec_->enqueue(r, msg);
}

void POpContext::notifyComplete() {
  if (complete_) return;
  complete_ = true;
  tell(std::make_shared<CompleteMessage>(op_->name()));
}

void POpContext::notifyError(const std::string& content) { exec_->reportError(op_->name(), content); }

void connect(const std::shared_ptr<PhysicalOp>& producer, const std::shared_ptr<PhysicalOp>& consumer) {
  producer->produce(consumer);
  consumer->consume(producer);
}

void Executor::put(const std::shared_ptr<PhysicalOp>& op) {
  if (ops_.count(op->name())) throw std::runtime_error("duplicate operator name '" + op->name() + "'");
  ops_[op->name()] = op;
  order_.push_back(op->name());
}

std::shared_ptr<PhysicalOp> Executor::get(const std::string& name) const {
  auto it = ops_.find(name);
  return it == ops_.end() ? nullptr : it->second;
}

void Executor::enqueue(const std::string& recipient, const std::shared_ptr<Message>& msg) {
  queue_.emplace_back(recipient, msg);
}

void Executor::reportError(const std::string& sender, const std::string& content) {
  if (!failed_) {
    failed_ = true;
    error_ = "Error in operator '" + sender + "': " + content;
  }
}

Expected<void> Executor::execute() {
  for (auto& name : order_) {
    auto& op = ops_[name];
    auto ctx = std::make_shared<POpContext>(op.get(), this);
    for (auto& p : op->producers()) ctx->operatorMap().addProducer(p);
    op->create(ctx);
  }
  for (auto& name : order_) enqueue(name, std::make_shared<StartMessage>("root"));
  while (!queue_.empty() && !failed_) {
    auto [recipient, msg] = queue_.front();
    queue_.pop_front();
    auto it = ops_.find(recipient);
    if (it == ops_.end()) {
      reportError(msg->sender(), "message to unknown operator '" + recipient + "'");
      break;
    }
    auto& op = it->second;
    if (msg->type() == MessageType::COMPLETE) op->ctx()->operatorMap().setComplete(msg->sender());
    auto t0 = std::chrono::steady_clock::now();
    try {
      op->onReceive(Envelope(msg));
    } catch (const std::exception& e) {
      reportError(recipient, e.what());
    }
    processingMs_[recipient] +=
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  if (failed_) return make_unexpected(error_);
  for (auto& name : order_)
    if (!ops_[name]->ctx()->isComplete())
      return make_unexpected("operator '" + name + "' did not complete");
  return {};
}

}  // namespace fpdb::amd

// Messages between physical operators (fpdb-executor/include/fpdb/executor/message/
// MessageType.h:10-35, TupleSetMessage.h:19, BloomFilterMessage.h:18,
// CompleteMessage, ErrorMessage), delivered in Envelopes.
#pragma once

#include <memory>
#include <string>

#include "BloomFilter.h"
#include "TupleSet.h"

namespace fpdb::amd {

enum class MessageType { START, TUPLESET, BLOOM_FILTER, COMPLETE, ERROR };

class Message {
 public:
  Message(MessageType type, std::string sender) : type_(type), sender_(std::move(sender)) {}
  virtual ~Message() = default;
  MessageType type() const { return type_; }
  const std::string& sender() const { return sender_; }
  std::string getTypeString() const;

 private:
  MessageType type_;
  std::string sender_;
};

class StartMessage : public Message {
 public:
  explicit StartMessage(std::string sender) : Message(MessageType::START, std::move(sender)) {}
};

class TupleSetMessage : public Message {
 public:
  TupleSetMessage(std::shared_ptr<TupleSet> tuples, std::string sender)
      : Message(MessageType::TUPLESET, std::move(sender)), tuples_(std::move(tuples)) {}
  const std::shared_ptr<TupleSet>& tuples() const { return tuples_; }

 private:
  std::shared_ptr<TupleSet> tuples_;
};

class BloomFilterMessage : public Message {
 public:
  BloomFilterMessage(std::shared_ptr<BloomFilterBase> bf, std::string sender)
      : Message(MessageType::BLOOM_FILTER, std::move(sender)), bloomFilter_(std::move(bf)) {}
  const std::shared_ptr<BloomFilterBase>& getBloomFilter() const { return bloomFilter_; }

 private:
  std::shared_ptr<BloomFilterBase> bloomFilter_;
};

class CompleteMessage : public Message {
 public:
  explicit CompleteMessage(std::string sender) : Message(MessageType::COMPLETE, std::move(sender)) {}
};

class ErrorMessage : public Message {
 public:
  ErrorMessage(std::string content, std::string sender)
      : Message(MessageType::ERROR, std::move(sender)), content_(std::move(content)) {}
  const std::string& getContent() const { return content_; }

 private:
  std::string content_;
};

class Envelope {
 public:
  explicit Envelope(std::shared_ptr<Message> m) : message_(std::move(m)) {}
  const Message& message() const { return *message_; }
  const std::shared_ptr<Message>& ptr() const { return message_; }

 private:
  std::shared_ptr<Message> message_;
};

}  // namespace fpdb::amd

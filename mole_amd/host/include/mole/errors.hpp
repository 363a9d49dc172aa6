// Errors and results of the host layer: the variants of errors::Error
// (src/errors/src/lib.rs:7-15) and Rust's Result<T, Error>, plus Panic for
// the reference's assert!/unimplemented!/expect aborts.
#pragma once

#include <stdexcept>
#include <string>
#include <utility>
#include <variant>

#include "mole_gpu.h"

namespace mole {

enum class ErrorKind {
  LinalgError,               // errors::Error::LinalgError
  ShapeError,                // errors::Error::ShapeError
  FuncError,                 // errors::Error::FuncError (device failures map here)
  OperatorValueAccessError,  // errors::Error::OperatorValueAccessError
  DataAccessError,           // errors::Error::DataAccessError
  EmptyCacheError,           // errors::Error::EmptyCacheError
};

struct Error {
  ErrorKind kind;
  std::string message;
};

inline const char *error_name(ErrorKind k) {
  switch (k) {
    case ErrorKind::LinalgError: return "LinalgError";
    case ErrorKind::ShapeError: return "ShapeError";
    case ErrorKind::FuncError: return "FuncError";
    case ErrorKind::OperatorValueAccessError: return "OperatorValueAccessError";
    case ErrorKind::DataAccessError: return "DataAccessError";
    default: return "EmptyCacheError";
  }
}

// A Rust panic (assert!, unimplemented!, expect): not a recoverable Error.
struct Panic : std::logic_error {
  using std::logic_error::logic_error;
};

template <typename T>
class Result {
 public:
  Result(T v) : v_(std::move(v)) {}
  Result(Error e) : v_(std::move(e)) {}
  bool is_ok() const { return v_.index() == 0; }
  bool is_err() const { return !is_ok(); }
  const T &value() const {
    if (!is_ok()) throw Panic(std::string("called unwrap on an Err: ") + error().message);
    return std::get<0>(v_);
  }
  T &value() {
    if (!is_ok()) throw Panic(std::string("called unwrap on an Err: ") + error().message);
    return std::get<0>(v_);
  }
  const Error &error() const { return std::get<1>(v_); }

 private:
  std::variant<T, Error> v_;
};

struct Unit {};
using Status = Result<Unit>;
inline Status Ok() { return Status(Unit{}); }

// mole_status -> Error: codes 1..6 are the errors::Error variants in
// declaration order; device, unsupported, invalid-argument and
// out-of-memory failures become FuncError (the reference's closed enum has
// no device variant).
inline Error error_from_status(mole_status s) {
  const std::string msg = std::string(mole_status_string(s)) + ": " + mole_last_error();
  switch (s) {
    case MOLE_ERR_LINALG: return {ErrorKind::LinalgError, msg};
    case MOLE_ERR_SHAPE: return {ErrorKind::ShapeError, msg};
    case MOLE_ERR_FUNC: return {ErrorKind::FuncError, msg};
    case MOLE_ERR_OPERATOR_VALUE_ACCESS: return {ErrorKind::OperatorValueAccessError, msg};
    case MOLE_ERR_DATA_ACCESS: return {ErrorKind::DataAccessError, msg};
    case MOLE_ERR_EMPTY_CACHE: return {ErrorKind::EmptyCacheError, msg};
    default: return {ErrorKind::FuncError, msg};
  }
}

// The status of a C-ABI call as a Result: MOLE_ERR_INVALID_ARGUMENT is what
// the reference answers with a panic (an assert), so it throws Panic.
inline Status status(mole_status s) {
  if (s == MOLE_OK) return Ok();
  if (s == MOLE_ERR_INVALID_ARGUMENT)
    throw Panic(std::string(mole_status_string(s)) + ": " + mole_last_error());
  return error_from_status(s);
}

}  // namespace mole

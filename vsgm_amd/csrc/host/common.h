// Shared scalar types and small enums of the VSGM host side.
//
// Types follow the reference's src/common/meta.h:9-12 (u32 vertex ids,
// u64 edge offsets, u64 counts) and the condition / query enums of
// src/common/meta.h:35-44.
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace vsgm {

typedef uint32_t uintV;  // vertex id
typedef uint64_t uintE;  // edge offset
typedef uint64_t uintC;  // embedding count

constexpr int kMaxQueryVerticesNum = 7;  // meta.h:33
constexpr uintV kMaxuintV = 0xffffffffu;

// Numeric values match the reference enum (meta.h:35) and the C-ABI.
enum CondOperator { LESS_THAN = 0, LARGER_THAN = 1, NON_EQUAL = 2, OPERATOR_NONE = 3 };
enum QueryType { PatternType = 0, CliqueType = 1, GraphletType = 2 };

typedef std::vector<uintV> ConnType;
typedef std::vector<ConnType> AllConnType;
typedef std::pair<CondOperator, uintV> CondType;
typedef std::vector<std::vector<CondType>> AllCondType;

struct Error : public std::runtime_error {
  explicit Error(const std::string& m) : std::runtime_error(m) {}
};

#define VSGM_CHECK(cond, msg)                                                   \
  do {                                                                          \
    if (!(cond)) throw ::vsgm::Error(std::string(__FILE__) + ":" +              \
                                     std::to_string(__LINE__) + ": " + (msg)); \
  } while (0)

}  // namespace vsgm

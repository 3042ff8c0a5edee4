// Strict JSON for the wire contract: the request, artifact and batch documents.
//
// The grammar is RFC 8259 with the reference's restrictions (io/strict_json.cpp:15-240):
// duplicate object keys, trailing data, non-finite numbers and \u escapes are rejected,
// and every number keeps its source text so the canonical identity hashes the original
// spelling.  Object members are held sorted by key bytes, which is also the canonical
// member order (io/strict_json.cpp:448-532).
#pragma once

#include <cstddef>
#include <cstdint>
#include <initializer_list>
#include <map>
#include <string>
#include <vector>

namespace femm
{
namespace host
{

struct Json {
  enum class Type { kNull, kBool, kNumber, kString, kArray, kObject };
  Type type = Type::kNull;
  bool boolean = false;
  double number = 0.0;
  std::string text;                    // string value, or a number's source text
  std::vector<Json> items;             // array elements
  std::map<std::string, Json> members; // object members, sorted by key bytes
};

// Parses the whole of `text` (surrounding whitespace allowed).  On failure returns false
// and sets *error.
bool ParseJson(const std::string &text, Json *out, std::string *error);

// `value` is an object holding exactly these member names.
bool ExactMembers(const Json &value, std::initializer_list<const char *> names, std::string *error);

// The member `name` of object `value` if present with type `type`, else nullptr and *error.
const Json *Member(const Json &value, const char *name, Json::Type type, std::string *error);

// An integral number within int32 range.
bool Int32Of(const Json &value, int32_t *out);

// 64 hex digits; `lower_only` admits only 0-9 and a-f.
bool IsSha256Hex(const std::string &text, bool lower_only);

// Every element of an array is a number: their values.
bool NumberArray(const Json &value, std::vector<double> *out);

// SHA-256 of a byte string, as 64 lower-case hex digits (FIPS 180-4).
std::string Sha256Hex(const std::string &bytes);

// SHA-256 of the canonical text of a value: members in key-byte order, no whitespace,
// numbers in their source spelling, strings re-escaped minimally.
std::string CanonicalSha256(const Json &value);

// `text` escaped for a JSON string body.
std::string EscapeJson(const std::string &text);

}  // namespace host
}  // namespace femm

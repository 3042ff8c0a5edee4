// Strict JSON parser, canonical form and SHA-256 of the wire contract.

#include "json.hpp"

#include <charconv>
#include <cmath>
#include <cpuid.h>
#include <cstdlib>
#include <immintrin.h>
#include <cstring>
#include <limits>

namespace femm
{
namespace host
{
namespace
{

constexpr int kMaxDepth = 256;   // nesting limit: a hostile document cannot exhaust the stack

class Parser
{
public:
  explicit Parser(const std::string &text) : s_(text) {}

  bool Document(Json *out, std::string *error)
  {
    Space();
    if (!Value(out, 0, error))
      return false;
    Space();
    if (at_ != s_.size())
      return Fail(error, "trailing JSON data");
    return true;
  }

private:
  const std::string &s_;
  size_t at_ = 0;

  static bool Fail(std::string *error, const char *what)
  {
    *error = what;
    return false;
  }

  void Space()
  {
    while (at_ < s_.size() &&
           (s_[at_] == ' ' || s_[at_] == '\t' || s_[at_] == '\n' || s_[at_] == '\r'))
      ++at_;
  }

  bool Literal(const char *word)
  {
    const size_t n = std::strlen(word);
    if (s_.compare(at_, n, word) != 0)
      return false;
    at_ += n;
    return true;
  }

  bool Value(Json *out, int depth, std::string *error)
  {
    if (depth > kMaxDepth)
      return Fail(error, "JSON nesting too deep");
    if (at_ >= s_.size())
      return Fail(error, "unexpected JSON end");
    const char c = s_[at_];
    if (c == '{')
      return Object(out, depth, error);
    if (c == '[')
      return Array(out, depth, error);
    if (c == '"') {
      out->type = Json::Type::kString;
      return String(&out->text, error);
    }
    if (Literal("true")) {
      out->type = Json::Type::kBool;
      out->boolean = true;
      return true;
    }
    if (Literal("false")) {
      out->type = Json::Type::kBool;
      out->boolean = false;
      return true;
    }
    if (Literal("null")) {
      out->type = Json::Type::kNull;
      return true;
    }
    return Number(out, error);
  }

  bool String(std::string *out, std::string *error)
  {
    ++at_;   // opening quote
    out->clear();
    while (at_ < s_.size()) {
      const unsigned char c = static_cast<unsigned char>(s_[at_++]);
      if (c == '"')
        return true;
      if (c < 0x20)
        return Fail(error, "control character in JSON string");
      if (c != '\\') {
        out->push_back(static_cast<char>(c));
        continue;
      }
      if (at_ >= s_.size())
        return Fail(error, "truncated JSON escape");
      switch (s_[at_++]) {
      case '"': out->push_back('"'); break;
      case '\\': out->push_back('\\'); break;
      case '/': out->push_back('/'); break;
      case 'b': out->push_back('\b'); break;
      case 'f': out->push_back('\f'); break;
      case 'n': out->push_back('\n'); break;
      case 'r': out->push_back('\r'); break;
      case 't': out->push_back('\t'); break;
      default:
        return Fail(error, "unsupported JSON string escape");
      }
    }
    return Fail(error, "unterminated JSON string");
  }

  static bool Digit(char c) { return c >= '0' && c <= '9'; }

  bool Number(Json *out, std::string *error)
  {
    const size_t begin = at_;
    if (at_ < s_.size() && s_[at_] == '-')
      ++at_;
    if (at_ >= s_.size() || !Digit(s_[at_]))
      return Fail(error, "invalid JSON value");
    if (s_[at_] == '0') {
      ++at_;
    } else {
      while (at_ < s_.size() && Digit(s_[at_]))
        ++at_;
    }
    if (at_ < s_.size() && s_[at_] == '.') {
      ++at_;
      const size_t digits = at_;
      while (at_ < s_.size() && Digit(s_[at_]))
        ++at_;
      if (at_ == digits)
        return Fail(error, "invalid JSON fraction");
    }
    if (at_ < s_.size() && (s_[at_] == 'e' || s_[at_] == 'E')) {
      ++at_;
      if (at_ < s_.size() && (s_[at_] == '+' || s_[at_] == '-'))
        ++at_;
      const size_t digits = at_;
      while (at_ < s_.size() && Digit(s_[at_]))
        ++at_;
      if (at_ == digits)
        return Fail(error, "invalid JSON exponent");
    }
    out->type = Json::Type::kNumber;
    out->text.assign(s_, begin, at_ - begin);
    // Correctly rounded, as strtod, without the locale; out of range (overflow, or an
    // underflow that strtod takes to a subnormal or zero) goes to strtod for its result.
    const auto r = std::from_chars(s_.data() + begin, s_.data() + at_, out->number);
    if (r.ptr != s_.data() + at_)
      return Fail(error, "JSON number conversion failed");
    if (r.ec == std::errc::result_out_of_range) {
      char *end = nullptr;
      out->number = std::strtod(out->text.c_str(), &end);
      if (end != out->text.c_str() + out->text.size())
        return Fail(error, "JSON number conversion failed");
    } else if (r.ec != std::errc()) {
      return Fail(error, "JSON number conversion failed");
    }
    if (!std::isfinite(out->number))
      return Fail(error, "nonfinite JSON number");
    return true;
  }

  bool Array(Json *out, int depth, std::string *error)
  {
    ++at_;
    out->type = Json::Type::kArray;
    out->items.clear();
    Space();
    if (at_ < s_.size() && s_[at_] == ']') {
      ++at_;
      return true;
    }
    for (;;) {
      Json item;
      if (!Value(&item, depth + 1, error))
        return false;
      out->items.push_back(std::move(item));
      Space();
      if (at_ >= s_.size())
        return Fail(error, "unterminated JSON array");
      const char c = s_[at_++];
      if (c == ']')
        return true;
      if (c != ',')
        return Fail(error, "JSON array comma expected");
      Space();
    }
  }

  bool Object(Json *out, int depth, std::string *error)
  {
    ++at_;
    out->type = Json::Type::kObject;
    out->members.clear();
    Space();
    if (at_ < s_.size() && s_[at_] == '}') {
      ++at_;
      return true;
    }
    for (;;) {
      if (at_ >= s_.size() || s_[at_] != '"')
        return Fail(error, "JSON object key expected");
      std::string key;
      if (!String(&key, error))
        return false;
      Space();
      if (at_ >= s_.size() || s_[at_++] != ':')
        return Fail(error, "JSON colon expected");
      Space();
      Json value;
      if (!Value(&value, depth + 1, error))
        return false;
      if (!out->members.emplace(std::move(key), std::move(value)).second)
        return Fail(error, "duplicate JSON object key");
      Space();
      if (at_ >= s_.size())
        return Fail(error, "unterminated JSON object");
      const char c = s_[at_++];
      if (c == '}')
        return true;
      if (c != ',')
        return Fail(error, "JSON object comma expected");
      Space();
    }
  }
};

// ---- SHA-256 (FIPS 180-4) ----------------------------------------------------------------

constexpr uint32_t kRound[64] = {
    0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u, 0x923f82a4u,
    0xab1c5ed5u, 0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u, 0x72be5d74u, 0x80deb1feu,
    0x9bdc06a7u, 0xc19bf174u, 0xe49b69c1u, 0xefbe4786u, 0x0fc19dc6u, 0x240ca1ccu, 0x2de92c6fu,
    0x4a7484aau, 0x5cb0a9dcu, 0x76f988dau, 0x983e5152u, 0xa831c66du, 0xb00327c8u, 0xbf597fc7u,
    0xc6e00bf3u, 0xd5a79147u, 0x06ca6351u, 0x14292967u, 0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu,
    0x53380d13u, 0x650a7354u, 0x766a0abbu, 0x81c2c92eu, 0x92722c85u, 0xa2bfe8a1u, 0xa81a664bu,
    0xc24b8b70u, 0xc76c51a3u, 0xd192e819u, 0xd6990624u, 0xf40e3585u, 0x106aa070u, 0x19a4c116u,
    0x1e376c08u, 0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au, 0x5b9cca4fu, 0x682e6ff3u,
    0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u, 0x90befffau, 0xa4506cebu, 0xbef9a3f7u,
    0xc67178f2u};

inline uint32_t Rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

void Compress(uint32_t h[8], const unsigned char *block)
{
  uint32_t w[64];
  for (int t = 0; t < 16; ++t)
    w[t] = static_cast<uint32_t>(block[4 * t]) << 24 | static_cast<uint32_t>(block[4 * t + 1]) << 16 |
           static_cast<uint32_t>(block[4 * t + 2]) << 8 | static_cast<uint32_t>(block[4 * t + 3]);
  for (int t = 16; t < 64; ++t) {
    const uint32_t s0 = Rotr(w[t - 15], 7) ^ Rotr(w[t - 15], 18) ^ (w[t - 15] >> 3);
    const uint32_t s1 = Rotr(w[t - 2], 17) ^ Rotr(w[t - 2], 19) ^ (w[t - 2] >> 10);
    w[t] = w[t - 16] + s0 + w[t - 7] + s1;
  }
  uint32_t v[8];
  std::memcpy(v, h, sizeof(v));
  for (int t = 0; t < 64; ++t) {
    const uint32_t e = v[4], a = v[0];
    const uint32_t t1 = v[7] + (Rotr(e, 6) ^ Rotr(e, 11) ^ Rotr(e, 25)) + ((e & v[5]) ^ (~e & v[6])) +
                        kRound[t] + w[t];
    const uint32_t t2 = (Rotr(a, 2) ^ Rotr(a, 13) ^ Rotr(a, 22)) + ((a & v[1]) ^ (a & v[2]) ^ (v[1] & v[2]));
    for (int k = 7; k > 0; --k)
      v[k] = v[k - 1];
    v[4] += t1;
    v[0] = t1 + t2;
  }
  for (int k = 0; k < 8; ++k)
    h[k] += v[k];
}

// The same compression with the x86 SHA extensions (the round and message-schedule
// instructions of FIPS 180-4's SHA-256; state kept as the ABEF / CDGH register pairs those
// instructions take).  Four rounds per step: W + K for rounds t..t+3, two rounds from the low
// half, two from the high half.
__attribute__((target("sha,sse4.1,ssse3"))) void CompressShaNi(uint32_t h[8],
                                                              const unsigned char *p,
                                                              size_t blocks)
{
  const __m128i order = _mm_set_epi64x(0x0c0d0e0f08090a0bLL, 0x0405060700010203LL);
  const __m128i dcba = _mm_loadu_si128(reinterpret_cast<const __m128i *>(h));
  const __m128i hgfe = _mm_loadu_si128(reinterpret_cast<const __m128i *>(h + 4));
  const __m128i badc = _mm_shuffle_epi32(dcba, 0xB1);
  const __m128i efgh = _mm_shuffle_epi32(hgfe, 0x1B);
  __m128i abef = _mm_alignr_epi8(badc, efgh, 8);
  __m128i cdgh = _mm_blend_epi16(efgh, badc, 0xF0);
  for (size_t b = 0; b < blocks; ++b, p += 64) {
    const __m128i abef0 = abef, cdgh0 = cdgh;
    __m128i w[4];
    for (int g = 0; g < 16; ++g) {
      __m128i m;
      if (g < 4) {
        m = _mm_shuffle_epi8(_mm_loadu_si128(reinterpret_cast<const __m128i *>(p + 16 * g)),
                             order);
      } else {
        // W[t] = s1(W[t-2]) + W[t-7] + s0(W[t-15]) + W[t-16] for t = 4g .. 4g+3.
        const __m128i x = _mm_add_epi32(_mm_sha256msg1_epu32(w[g & 3], w[(g + 1) & 3]),
                                        _mm_alignr_epi8(w[(g + 3) & 3], w[(g + 2) & 3], 4));
        m = _mm_sha256msg2_epu32(x, w[(g + 3) & 3]);
      }
      w[g & 3] = m;
      const __m128i wk =
          _mm_add_epi32(m, _mm_loadu_si128(reinterpret_cast<const __m128i *>(kRound + 4 * g)));
      // Two rounds give the new ABEF; the old ABEF is the new CDGH.  After the second pair
      // both variables hold their own roles again.
      cdgh = _mm_sha256rnds2_epu32(cdgh, abef, wk);
      abef = _mm_sha256rnds2_epu32(abef, cdgh, _mm_shuffle_epi32(wk, 0x0E));
    }
    abef = _mm_add_epi32(abef, abef0);
    cdgh = _mm_add_epi32(cdgh, cdgh0);
  }
  const __m128i feba = _mm_shuffle_epi32(abef, 0x1B);
  const __m128i ghcd = _mm_shuffle_epi32(cdgh, 0xB1);
  _mm_storeu_si128(reinterpret_cast<__m128i *>(h), _mm_blend_epi16(feba, ghcd, 0xF0));
  _mm_storeu_si128(reinterpret_cast<__m128i *>(h + 4), _mm_alignr_epi8(ghcd, feba, 8));
}

// SHA extensions present (CPUID leaf 7 EBX bit 29, SSSE3 and SSE4.1), unless
// FEMM_GPU_AMD_PORTABLE_SHA is set to a value other than 0.
bool UseShaNi()
{
  static const bool use = [] {
    const char *force = std::getenv("FEMM_GPU_AMD_PORTABLE_SHA");
    if (force != nullptr && *force != '\0' && std::strcmp(force, "0") != 0)
      return false;
    unsigned a = 0, b = 0, c = 0, d = 0;
    if (!__get_cpuid(1, &a, &b, &c, &d) || !(c & (1u << 9)) || !(c & (1u << 19)))
      return false;
    if (!__get_cpuid_count(7, 0, &a, &b, &c, &d))
      return false;
    return (b & (1u << 29)) != 0;
  }();
  return use;
}

void CompressBlocks(uint32_t h[8], const unsigned char *p, size_t blocks)
{
  if (UseShaNi()) {
    CompressShaNi(h, p, blocks);
    return;
  }
  for (size_t b = 0; b < blocks; ++b)
    Compress(h, p + 64 * b);
}

void AppendCanonical(const Json &v, std::string *out)
{
  switch (v.type) {
  case Json::Type::kNull:
    out->append("null");
    return;
  case Json::Type::kBool:
    out->append(v.boolean ? "true" : "false");
    return;
  case Json::Type::kNumber:
    out->append(v.text);
    return;
  case Json::Type::kString:
    out->push_back('"');
    out->append(EscapeJson(v.text));
    out->push_back('"');
    return;
  case Json::Type::kArray:
    out->push_back('[');
    for (size_t i = 0; i < v.items.size(); ++i) {
      if (i)
        out->push_back(',');
      AppendCanonical(v.items[i], out);
    }
    out->push_back(']');
    return;
  case Json::Type::kObject: {
    out->push_back('{');
    bool first = true;
    for (const auto &m : v.members) {
      if (!first)
        out->push_back(',');
      first = false;
      out->push_back('"');
      out->append(EscapeJson(m.first));
      out->append("\":");
      AppendCanonical(m.second, out);
    }
    out->push_back('}');
    return;
  }
  }
}

}  // namespace

bool ParseJson(const std::string &text, Json *out, std::string *error)
{
  *out = Json{};
  return Parser(text).Document(out, error);
}

bool ExactMembers(const Json &value, std::initializer_list<const char *> names, std::string *error)
{
  if (value.type != Json::Type::kObject || value.members.size() != names.size()) {
    *error = "unexpected JSON object fields";
    return false;
  }
  for (const char *name : names)
    if (value.members.find(name) == value.members.end()) {
      *error = std::string("missing JSON field: ") + name;
      return false;
    }
  return true;
}

const Json *Member(const Json &value, const char *name, Json::Type type, std::string *error)
{
  if (value.type != Json::Type::kObject) {
    *error = "expected JSON object";
    return nullptr;
  }
  const auto found = value.members.find(name);
  if (found == value.members.end()) {
    *error = std::string("missing JSON field: ") + name;
    return nullptr;
  }
  if (found->second.type != type) {
    *error = std::string("wrong JSON type: ") + name;
    return nullptr;
  }
  return &found->second;
}

bool Int32Of(const Json &value, int32_t *out)
{
  if (value.type != Json::Type::kNumber || value.number != std::floor(value.number) ||
      value.number < static_cast<double>(std::numeric_limits<int32_t>::min()) ||
      value.number > static_cast<double>(std::numeric_limits<int32_t>::max()))
    return false;
  *out = static_cast<int32_t>(value.number);
  return true;
}

bool IsSha256Hex(const std::string &text, bool lower_only)
{
  if (text.size() != 64)
    return false;
  for (const char c : text) {
    const bool digit = c >= '0' && c <= '9', lower = c >= 'a' && c <= 'f';
    const bool upper = c >= 'A' && c <= 'F';
    if (!(digit || lower || (!lower_only && upper)))
      return false;
  }
  return true;
}

bool NumberArray(const Json &value, std::vector<double> *out)
{
  if (value.type != Json::Type::kArray)
    return false;
  out->clear();
  for (const Json &item : value.items) {
    if (item.type != Json::Type::kNumber)
      return false;
    out->push_back(item.number);
  }
  return true;
}

std::string Sha256Hex(const std::string &bytes)
{
  uint32_t h[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au,
                   0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};
  const size_t n = bytes.size();
  const unsigned char *p = reinterpret_cast<const unsigned char *>(bytes.data());
  const size_t full = n / 64;
  CompressBlocks(h, p, full);
  // Padding: the remaining bytes, 0x80, zeros, the bit length (big-endian) in 64 bits.
  unsigned char tail[128] = {};
  const size_t rest = n - 64 * full;
  std::memcpy(tail, p + 64 * full, rest);
  tail[rest] = 0x80;
  const size_t blocks = rest + 9 <= 64 ? 1 : 2;
  const uint64_t bits = static_cast<uint64_t>(n) * 8u;
  for (int k = 0; k < 8; ++k)
    tail[64 * blocks - 1 - k] = static_cast<unsigned char>(bits >> (8 * k));
  CompressBlocks(h, tail, blocks);
  static const char *digits = "0123456789abcdef";
  std::string hex(64, '0');
  for (int k = 0; k < 8; ++k)
    for (int j = 0; j < 8; ++j)
      hex[8 * k + j] = digits[(h[k] >> (28 - 4 * j)) & 0xfu];
  return hex;
}

std::string CanonicalSha256(const Json &value)
{
  std::string text;
  AppendCanonical(value, &text);
  return Sha256Hex(text);
}

std::string EscapeJson(const std::string &text)
{
  static const char *digits = "0123456789abcdef";
  std::string out;
  out.reserve(text.size());
  for (const char ch : text) {
    const unsigned char c = static_cast<unsigned char>(ch);
    switch (c) {
    case '"': out.append("\\\""); break;
    case '\\': out.append("\\\\"); break;
    case '\b': out.append("\\b"); break;
    case '\f': out.append("\\f"); break;
    case '\n': out.append("\\n"); break;
    case '\r': out.append("\\r"); break;
    case '\t': out.append("\\t"); break;
    default:
      if (c < 0x20) {
        out.append("\\u00");
        out.push_back(digits[c >> 4]);
        out.push_back(digits[c & 0xf]);
      } else {
        out.push_back(static_cast<char>(c));
      }
    }
  }
  return out;
}

}  // namespace host
}  // namespace femm

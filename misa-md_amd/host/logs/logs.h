// The logging calls of the kiwi framework that the reference makes
// (kiwi::logs::e/w/i/s/v/d with a tag and a "{}"-placeholder format, some
// with a rank first: logs::i(MASTER_PROCESSOR, "sim", "...{}...", x)), for
// one rank: every call writes "[level] tag: message" to the sink, standard
// error by default.
#pragma once
#include <cstdio>
#include <sstream>
#include <string>
#include <type_traits>
#include <utility>

namespace kiwi {
namespace logs {

inline FILE*& sink() {
  static FILE* f = stderr;
  return f;
}
inline void setSink(FILE* f) { sink() = f ? f : stderr; }

namespace detail {
inline void format(std::ostringstream& os, const char* f) { os << f; }
template <typename A, typename... R>
void format(std::ostringstream& os, const char* f, A&& a, R&&... rest) {
  for (; *f; ++f) {
    if (f[0] == '{' && f[1] == '}') {
      os << a;
      format(os, f + 2, std::forward<R>(rest)...);
      return;
    }
    os << *f;
  }
}
template <typename... A>
void emit(char level, const std::string& tag, const char* fmt, A&&... a) {
  std::ostringstream os;
  format(os, fmt, std::forward<A>(a)...);
  std::fprintf(sink(), "[%c] %s: %s", level, tag.c_str(), os.str().c_str());
  std::fflush(sink());
}
}  // namespace detail

// one rank: the rank argument (the rank that should print) is always this one
// (the tag overload takes no integer first argument, so a rank of 0 is a rank)
#define MISA_KIWI_LOG_LEVEL(name, ch)                                                                 \
  template <typename Tag, typename... A,                                                              \
            typename = std::enable_if_t<!std::is_integral<std::decay_t<Tag>>::value>>                  \
  void name(Tag&& tag, const char* fmt, A&&... a) {                                                   \
    detail::emit(ch, std::string(tag), fmt, std::forward<A>(a)...);                                   \
  }                                                                                                   \
  template <typename... A>                                                           \
  void name(int rank, const std::string& tag, const char* fmt, A&&... a) {           \
    if (rank == 0) detail::emit(ch, tag, fmt, std::forward<A>(a)...);                \
  }
MISA_KIWI_LOG_LEVEL(e, 'E')
MISA_KIWI_LOG_LEVEL(w, 'W')
MISA_KIWI_LOG_LEVEL(i, 'I')
MISA_KIWI_LOG_LEVEL(s, 'S')
MISA_KIWI_LOG_LEVEL(v, 'V')
MISA_KIWI_LOG_LEVEL(d, 'D')
#undef MISA_KIWI_LOG_LEVEL

}  // namespace logs
}  // namespace kiwi

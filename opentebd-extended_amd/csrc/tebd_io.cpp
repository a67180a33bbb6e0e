// Checkpoint files of a Vidal state in the reference's own text formats, so that a state
// written here restarts the reference and the reference's files restart this library:
//   lambdas: RecordLambdas (ToolsCore_v_2.0/io_module.f90:191-219): L+1 records, one per
//            bond, of chi values;
//   gammas:  RecordGammas (io_module.f90:221-260): for every site l, for alpha = 1..chi,
//            for beta = 1..chi one record of the d real parts Re Gamma_l(alpha, i, beta),
//            i = 1..d; then the same records of the imaginary parts;
// read back as readGammaLambda does (io_module.f90:284-373).  Kind 'S' writes every value
// as Fortran's E30.15 edit descriptor does (sign, "0.", 15 significant digits, E and a
// signed two-digit exponent -- or a signed three-digit exponent without the E beyond
// +-99 -- right-justified in 30 columns); kind 'B' writes the value's 64 bits as B80 does
// (binary digits without leading zeros, right-justified in 80 columns), which restarts
// bit for bit.  Host-only: no device call.
#include "../../include/tebd_hip.h"

#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace {

// Fortran E30.15 of v.
std::string fmt_e30(double v) {
  char body[64];
  const bool neg = std::signbit(v);
  const double a = std::fabs(v);
  int exp10 = 0;
  std::string digits;
  if (a == 0.0) {
    digits = std::string(15, '0');
  } else {
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.14e", a);  // d.ddddddddddddddde+XX, correctly rounded
    digits.push_back(buf[0]);
    digits.append(buf + 2, 14);
    exp10 = std::atoi(std::strchr(buf, 'e') + 1) + 1;  // 0.d... form
  }
  if (std::abs(exp10) <= 99)
    std::snprintf(body, sizeof(body), "%s0.%sE%c%02d", neg ? "-" : "", digits.c_str(), exp10 < 0 ? '-' : '+',
                  std::abs(exp10));
  else
    std::snprintf(body, sizeof(body), "%s0.%s%c%03d", neg ? "-" : "", digits.c_str(), exp10 < 0 ? '-' : '+',
                  std::abs(exp10));
  std::string s(body);
  return std::string(s.size() < 30 ? 30 - s.size() : 0, ' ') + s;
}

// Fortran B80 of v's bit pattern.
std::string fmt_b80(double v) {
  uint64_t u;
  std::memcpy(&u, &v, sizeof(u));
  std::string s;
  if (u == 0) s = "0";
  while (u) {
    s.insert(s.begin(), (char)('0' + (u & 1u)));
    u >>= 1;
  }
  return std::string(80 - s.size(), ' ') + s;
}

bool parse_value(const std::string& tok, char kind, double* out) {
  if (kind == 'B') {
    uint64_t u = 0;
    if (tok.empty() || tok.size() > 64) return false;
    for (char ch : tok) {
      if (ch != '0' && ch != '1') return false;
      u = (u << 1) | (uint64_t)(ch - '0');
    }
    std::memcpy(out, &u, sizeof(u));
    return true;
  }
  // E form, possibly with the E left out before a three-digit exponent ("0.1-100")
  std::string t = tok;
  const size_t e = t.find_first_of("EeDd");
  if (e != std::string::npos) {
    t[e] = 'E';
  } else {
    const size_t sgn = t.find_last_of("+-");
    if (sgn != std::string::npos && sgn > 0) t.insert(sgn, "E");
  }
  char* end = nullptr;
  *out = std::strtod(t.c_str(), &end);
  return end && *end == '\0';
}

struct Writer {
  std::FILE* f;
  char kind;
  void record(const double* v, int n, size_t stride) {
    std::string line;
    for (int k = 0; k < n; ++k) line += kind == 'B' ? fmt_b80(v[k * stride]) : fmt_e30(v[k * stride]);
    line.push_back('\n');
    std::fwrite(line.data(), 1, line.size(), f);
  }
};

// The first n values of the next record (line) of a checkpoint file.
bool read_record(std::ifstream& in, char kind, int n, double* out, size_t stride) {
  std::string line;
  if (!std::getline(in, line)) return false;
  std::istringstream ss(line);
  std::string tok;
  for (int k = 0; k < n; ++k) {
    if (!(ss >> tok) || !parse_value(tok, kind, out + k * stride)) return false;
  }
  return true;
}

}  // namespace

extern "C" {

int tebd_write_state_files(int L, int d, int chi, const double* gamma, const double* lambda, const char* lambda_path,
                           const char* gamma_path, char kind) {
  if (L < 1 || d < 1 || chi < 1 || !gamma || !lambda || !lambda_path || !gamma_path || (kind != 'S' && kind != 'B'))
    return TEBD_ERR_ARG;
  std::FILE* fl = std::fopen(lambda_path, "w");
  if (!fl) return TEBD_ERR_ARG;
  Writer wl{fl, kind};
  for (int l = 0; l <= L; ++l) wl.record(lambda + (size_t)l * chi, chi, 1);
  const bool ok_l = std::fclose(fl) == 0;
  std::FILE* fg = std::fopen(gamma_path, "w");
  if (!fg) return TEBD_ERR_ARG;
  Writer wg{fg, kind};
  const size_t site = (size_t)2 * chi * d * chi;  // interleaved complex, Gamma(a, i, b) at a + chi (i + d b)
  for (int part = 0; part < 2; ++part)
    for (int s = 0; s < L; ++s)
      for (int a = 0; a < chi; ++a)
        for (int b = 0; b < chi; ++b) wg.record(gamma + s * site + 2 * ((size_t)a + (size_t)chi * d * b) + part, d,
                                                (size_t)2 * chi);
  const bool ok_g = std::fclose(fg) == 0;
  return ok_l && ok_g ? TEBD_OK : TEBD_ERR_ARG;
}

int tebd_read_state_files(int L, int d, int chi, double* gamma, double* lambda, const char* lambda_path,
                          const char* gamma_path, char kind) {
  if (L < 1 || d < 1 || chi < 1 || !gamma || !lambda || !lambda_path || !gamma_path || (kind != 'S' && kind != 'B'))
    return TEBD_ERR_ARG;
  std::ifstream il(lambda_path), ig(gamma_path);
  if (!il || !ig) return TEBD_ERR_ARG;
  for (int l = 0; l <= L; ++l)
    if (!read_record(il, kind, chi, lambda + (size_t)l * chi, 1)) return TEBD_ERR_SHAPE;
  const size_t site = (size_t)2 * chi * d * chi;
  for (int part = 0; part < 2; ++part)
    for (int s = 0; s < L; ++s)
      for (int a = 0; a < chi; ++a)
        for (int b = 0; b < chi; ++b)
          if (!read_record(ig, kind, d, gamma + s * site + 2 * ((size_t)a + (size_t)chi * d * b) + part, (size_t)2 * chi))
            return TEBD_ERR_SHAPE;
  return TEBD_OK;
}

}  // extern "C"

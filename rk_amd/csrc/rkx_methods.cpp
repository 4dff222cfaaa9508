// Butcher tables of the seven methods the reference ships (src/methods.c:62-724),
// as exact rationals in a small text format parsed once into rkx_method.
//
// Format, one record per line:  c|b|bh <s values> ;  a<i> <i values>
// A value is an integer or "p/q"; each is converted as (double)p / (double)q,
// which is the correctly rounded quotient, identical to the reference's
// literal divisions.  Unlisted entries are zero -- including VERNER65
// A[6][5], which the reference leaves unassigned (intended 0).  DOPRI87 keeps
// the reference's A[6][3] = 77736583/692538347 (src/methods.c:630) for
// parity, although the published Prince-Dormand value is 77736538/692538347.
#include "rkx_internal.hpp"

#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>

namespace {

struct table_src {
  const char *name;
  int s, ord, has_emb;
  const char *text;
};

const table_src kTables[] = {
    {"HEUN2", 2, 2, 0,
     "c 0 1\n"
     "b 1/2 1/2\n"
     "a1 1\n"},
    {"RK4", 4, 4, 0,
     "c 0 0.5 0.5 1\n"
     "b 1/6 2/6 2/6 1/6\n"
     "a1 0.5\n"
     "a2 0 0.5\n"
     "a3 0 0 1\n"},
    {"SSPRK3", 3, 3, 0,
     "c 0 1 1/2\n"
     "b 1/6 1/6 2/3\n"
     "a1 1\n"
     "a2 1/4 1/4\n"},
    {"RKF23", 3, 2, 1,
     "c 0 1 1/2\n"
     "b 1/2 1/2 0\n"
     "bh 1/6 1/6 4/6\n"
     "a1 1\n"
     "a2 1/4 1/4\n"},
    {"DOPRI54", 7, 5, 1,
     "c 0 1/5 3/10 4/5 8/9 1 1\n"
     "b 35/384 0 500/1113 125/192 -2187/6784 11/84 0\n"
     "bh 5179/57600 0 7571/16695 393/640 -92097/339200 187/2100 1/40\n"
     "a1 1/5\n"
     "a2 3/40 9/40\n"
     "a3 44/45 -56/15 32/9\n"
     "a4 19372/6561 -25360/2187 64448/6561 -212/729\n"
     "a5 9017/3168 -355/33 46732/5247 49/176 -5103/18656\n"
     "a6 35/384 0 500/1113 125/192 -2187/6784 11/84\n"},
    {"VERNER65", 8, 6, 1,
     "c 0 1/18 1/6 2/9 2/3 1 8/9 1\n"
     "b 3/80 0 4/25 243/1120 77/160 73/700 0 0\n"
     "bh 57/640 0 -16/65 1377/2240 121/320 0 891/8320 2/35\n"
     "a1 1/18\n"
     "a2 -1/12 1/4\n"
     "a3 -2/81 4/27 8/81\n"
     "a4 40/33 -4/11 -56/11 54/11\n"
     "a5 -369/73 72/73 5380/219 -12285/584 2695/1752\n"
     "a6 -8716/891 656/297 39520/891 -416/11 52/27 0\n"
     "a7 3015/256 -9/4 -4219/78 5985/128 -539/384 0 693/3328\n"},
    {"DOPRI87", 13, 8, 1,
     "c 0 1/18 1/12 1/8 5/16 3/8 59/400 93/200 5490023248/9719169821 13/20 "
     "1201146811/1299019798 1 1\n"
     "b 14005451/335480064 0 0 0 0 -59238493/1068277825 181606767/758867731 "
     "561292985/797845732 -1041891430/1371343529 760417239/1151165299 "
     "118820643/751138087 -528747749/2220607170 1/4\n"
     "bh 13451932/455176623 0 0 0 0 -808719846/976000145 1757004468/5645159321 "
     "656045339/265891186 -3867574721/1518517206 465885868/322736535 "
     "53011238/667516719 2/45 0\n"
     "a1 1/18\n"
     "a2 1/48 1/16\n"
     "a3 1/32 0 3/32\n"
     "a4 5/16 0 -75/64 75/64\n"
     "a5 3/80 0 0 3/16 3/20\n"
     "a6 29443841/614563906 0 0 77736583/692538347 -28693883/1125000000 "
     "23124283/1800000000\n"
     "a7 16016141/946692911 0 0 61564180/158732637 22789713/633445777 "
     "545815736/2771057229 -180193667/1043307555\n"
     "a8 39632708/573591083 0 0 -433636366/683701615 -421739975/2616292301 "
     "100302831/723423059 790204164/839813087 800635310/3783071287\n"
     "a9 246121993/1340847787 0 0 -37695042795/15268766246 -309121744/1061227803 "
     "-12992083/490766935 6005943493/2108947869 393006217/1396673457 "
     "123872331/1001029789\n"
     "a10 -1028468189/846180014 0 0 8478235783/508512852 1311729495/1432422823 "
     "-10304129995/1701304382 -48777925059/3047939560 15336726248/1032824649 "
     "-45442868181/3398467696 3065993473/597172653\n"
     "a11 185892177/718116043 0 0 -3185094517/667107341 -477755414/1098053517 "
     "-703635378/230739211 5731566787/1027545527 5232866602/850066563 "
     "-4093664535/808688257 3962137247/1805957418 65686358/487910083\n"
     "a12 403863854/491063109 0 0 -5068492393/434740067 -411421997/543043805 "
     "652783627/914296604 11173962825/925320556 -13158990841/6184727034 "
     "3936647629/1978049680 -160528059/685178525 248638103/1413531060 0\n"},
};

double parse_rational(const std::string &tok)
{
  const size_t slash = tok.find('/');
  if (slash == std::string::npos)
    return std::strtod(tok.c_str(), nullptr);
  const double p = std::strtod(tok.substr(0, slash).c_str(), nullptr);
  const double q = std::strtod(tok.substr(slash + 1).c_str(), nullptr);
  return p / q;
}

bool parse_table(const table_src &src, rkx_method *m)
{
  std::memset(m, 0, sizeof(*m));
  std::strncpy(m->name, src.name, sizeof(m->name) - 1);
  m->s = src.s;
  m->ord = src.ord;
  m->has_emb = src.has_emb;
  std::istringstream in(src.text);
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream ls(line);
    std::string key, tok;
    ls >> key;
    double *dst = nullptr;
    int cap = src.s;
    if (key == "c") dst = m->c;
    else if (key == "b") dst = m->b;
    else if (key == "bh") dst = m->b_hat;
    else if (key.size() > 1 && key[0] == 'a') {
      const int row = std::atoi(key.c_str() + 1);
      if (row < 1 || row >= src.s) return false;
      dst = m->A[row];
      cap = row;                       // explicit method: A[i][j] = 0 for j >= i
    } else return false;
    int k = 0;
    while (ls >> tok) {
      if (k >= cap) return false;
      dst[k++] = parse_rational(tok);
    }
  }
  return true;
}

}  // namespace

extern "C" int rkx_method_get(const char *name, rkx_method *m)
{
  if (!name || !m)
    return rkx::fail(RKX_EINVAL, "rkx_method_get: null argument");
  for (const auto &t : kTables)
    if (std::strcmp(t.name, name) == 0) {
      if (!parse_table(t, m))
        return rkx::fail(RKX_EINVAL, "rkx_method_get: malformed table");
      return RKX_OK;
    }
  return rkx::fail(RKX_EINVAL, std::string("rkx_method_get: unknown method ") + name);
}

// java_text.cpp -- Java's Float.toString and the OpenCL literal read-back of the constants
// LiveSRRF_CL bakes into its kernel text (LiveSRRF_CL.java:311-342).
#include "java_text.h"

#include <locale.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

namespace esrrf {

namespace {

// printf/strtod under the "C" numeric locale for the calling thread only, so a host
// process with a decimal comma does not change the digits.
locale_t c_numeric_locale()
{
    static locale_t loc = newlocale(LC_NUMERIC_MASK, "C", (locale_t)0);
    return loc;
}

struct ScopedCLocale {
    locale_t old = (locale_t)0;
    bool on = false;
    ScopedCLocale()
    {
        locale_t c = c_numeric_locale();
        if (c != (locale_t)0) {
            old = uselocale(c);
            on = true;
        }
    }
    ~ScopedCLocale()
    {
        if (on) uselocale(old);
    }
};

// value = digits * 10^e10
struct Dec {
    uint64_t digits;
    int e10;
};

void strip_zeros(Dec &d)
{
    while (d.digits != 0 && d.digits % 10 == 0) {
        d.digits /= 10;
        d.e10++;
    }
}

int n_digits(uint64_t v)
{
    int n = 0;
    do {
        n++;
        v /= 10;
    } while (v);
    return n;
}

float dec_float(const Dec &d)
{
    char buf[48];
    snprintf(buf, sizeof buf, "%llue%d", (unsigned long long)d.digits, d.e10);
    return strtof(buf, nullptr);
}

long double dec_long_double(const Dec &d)
{
    char buf[48];
    snprintf(buf, sizeof buf, "%llue%d", (unsigned long long)d.digits, d.e10);
    return strtold(buf, nullptr);
}

// The decimal Float.toString renders for a > 0 finite: among the decimals that round to a
// (round to nearest), those of minimal length m (m >= 2), or of length 1 or 2 when m == 1;
// of these the one closest to a, the even last digit on a tie.
Dec shortest_decimal(float a)
{
    std::vector<Dec> cand;
    for (int q = 1; q <= 9; q++) {
        char buf[64];
        snprintf(buf, sizeof buf, "%.*e", q - 1, (double)a);
        uint64_t D = 0;
        const char *c = buf;
        for (; *c && *c != 'e'; c++)
            if (*c >= '0' && *c <= '9') D = D * 10 + (uint64_t)(*c - '0');
        const int E = atoi(c + 1);
        const int e10 = E - (q - 1);
        cand.push_back({D, e10});
        cand.push_back({D + 1, e10});
        if (D > 1) cand.push_back({D - 1, e10});
        uint64_t p10 = 1;
        for (int k = 1; k < q; k++) p10 *= 10;
        if (D == p10) cand.push_back({p10 * 10 - 1, e10 - 1});   // the q-digit decimal below
    }
    int m = 100;
    std::vector<Dec> ok;
    for (Dec d : cand) {
        strip_zeros(d);
        if (d.digits == 0 || dec_float(d) != a) continue;
        ok.push_back(d);
        m = std::min(m, n_digits(d.digits));
    }
    const int mmax = m == 1 ? 2 : m;
    Dec best{0, 0};
    long double bestd = -1;
    for (const Dec &d : ok) {
        const int n = n_digits(d.digits);
        if (n < m || n > mmax) continue;
        const long double dist = std::fabs(dec_long_double(d) - (long double)a);
        const bool closer = bestd < 0 || dist < bestd;
        const bool tie_even = dist == bestd && (d.digits % 2 == 0) && (best.digits % 2 != 0);
        if (closer || tie_even) {
            best = d;
            bestd = dist;
        }
    }
    return best;
}

}  // namespace

std::string java_float_to_string(float x)
{
    if (std::isnan(x)) return "NaN";
    if (std::isinf(x)) return x > 0 ? "Infinity" : "-Infinity";
    if (x == 0.0f) return std::signbit(x) ? "-0.0" : "0.0";
    ScopedCLocale cl;
    const float a = std::fabs(x);
    const Dec d = shortest_decimal(a);
    const std::string s = std::to_string((unsigned long long)d.digits);
    const int n = (int)s.size();
    const int esci = d.e10 + n - 1;    // a = s[0].s[1..] * 10^esci
    std::string out = x < 0 ? "-" : "";
    if (esci >= -3 && esci <= 6) {    // 1e-3 <= decimal < 1e7
        if (esci >= 0) {
            std::string ip = s.substr(0, (size_t)std::min(n, esci + 1));
            while ((int)ip.size() < esci + 1) ip += '0';
            std::string fp = esci + 1 < n ? s.substr((size_t)esci + 1) : "0";
            out += ip + "." + fp;
        } else {
            out += "0." + std::string((size_t)(-esci - 1), '0') + s;
        }
    } else {
        out += s.substr(0, 1) + "." + (n > 1 ? s.substr(1) : "0") + "E" + std::to_string(esci);
    }
    return out;
}

float opencl_literal_to_float(const std::string &text)
{
    ScopedCLocale cl;
    const double v = strtod(text.c_str(), nullptr);
    return (float)v;
}

float java_two_sigma_plus_one(float two_sigma)
{
    return opencl_literal_to_float(java_float_to_string(two_sigma) + "1");
}

}  // namespace esrrf

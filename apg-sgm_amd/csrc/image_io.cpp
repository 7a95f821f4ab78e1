// PGM / PPM / PFM file I/O of the host API (formats of reference
// src/image.cpp:45-205).
#include <stdint.h>
#include <string.h>

#include <cctype>
#include <fstream>
#include <string>
#include <vector>

#include "apg_sgm/image.hpp"

namespace apg {

namespace {

// next whitespace-separated header token, '#' comments skipped to the end of line
bool header_token(std::istream& is, std::string& tok) {
    tok.clear();
    int c;
    for (;;) {
        c = is.get();
        if (c == EOF) return false;
        if (c == '#') {
            while ((c = is.get()) != EOF && c != '\n') {
            }
            continue;
        }
        if (!std::isspace(c)) break;
    }
    tok.push_back(static_cast<char>(c));
    while ((c = is.get()) != EOF && !std::isspace(c)) tok.push_back(static_cast<char>(c));
    return true;
}

bool parse_int(const std::string& s, int& v) {
    if (s.empty()) return false;
    try {
        size_t k = 0;
        v = std::stoi(s, &k);
        return k == s.size();
    } catch (...) {
        return false;
    }
}

}  // namespace

bool load_image(const std::string& path, Image8& out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::string magic, sw, sh, smax;
    int w, h, maxv;
    if (!header_token(f, magic) || !header_token(f, sw) || !header_token(f, sh) || !header_token(f, smax)) return false;
    if (!parse_int(sw, w) || !parse_int(sh, h) || !parse_int(smax, maxv)) return false;
    if (w <= 0 || h <= 0 || maxv <= 0 || maxv > 255) return false;
    const int ch = magic == "P5" ? 1 : magic == "P6" ? 3 : 0;
    if (!ch) return false;
    Image8 img(w, h, ch);
    f.read(reinterpret_cast<char*>(img.data()), static_cast<std::streamsize>(w) * h * ch);
    if (!f) return false;
    out = std::move(img);
    return true;
}

bool save_pgm(const std::string& path, const Image8& img) {
    if (img.empty() || img.channels() != 1) return false;
    std::ofstream f(path, std::ios::binary);
    if (!f) return false;
    f << "P5\n" << img.width() << " " << img.height() << "\n255\n";
    f.write(reinterpret_cast<const char*>(img.data()), static_cast<std::streamsize>(img.width()) * img.height());
    return static_cast<bool>(f);
}

bool load_pfm(const std::string& path, Image32f& out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::string tag, sw, sh, ss;
    if (!header_token(f, tag) || !header_token(f, sw) || !header_token(f, sh) || !header_token(f, ss)) return false;
    const int ch = tag == "Pf" ? 1 : tag == "PF" ? 3 : 0;
    int w, h;
    if (!ch || !parse_int(sw, w) || !parse_int(sh, h) || w <= 0 || h <= 0) return false;
    float scale;
    try {
        scale = std::stof(ss);
    } catch (...) {
        return false;
    }
    const bool big_endian = scale > 0.f;   // negative scale: little endian
    Image32f img(w, h, 0.f);
    std::vector<uint32_t> row(static_cast<size_t>(w) * ch);
    for (int y = h - 1; y >= 0; --y) {   // rows are stored bottom-up
        f.read(reinterpret_cast<char*>(row.data()), static_cast<std::streamsize>(row.size() * 4));
        if (!f) return false;
        for (int x = 0; x < w; ++x) {
            uint32_t u = row[static_cast<size_t>(x) * ch];
            if (big_endian) u = __builtin_bswap32(u);
            float v;
            memcpy(&v, &u, 4);
            img.at(x, y) = v;
        }
    }
    out = std::move(img);
    return true;
}

bool save_pfm(const std::string& path, const Image32f& img) {
    if (img.empty()) return false;
    std::ofstream f(path, std::ios::binary);
    if (!f) return false;
    f << "Pf\n" << img.width() << " " << img.height() << "\n-1.0\n";
    for (int y = img.height() - 1; y >= 0; --y)
        f.write(reinterpret_cast<const char*>(img.data() + static_cast<size_t>(y) * img.width()),
                static_cast<std::streamsize>(img.width()) * 4);
    return static_cast<bool>(f);
}

bool save_disparity_preview(const std::string& path, const Image32f& disp, float dmax) {
    if (disp.empty()) return false;
    const float k = dmax > 1.f ? 255.f / dmax : 1.f;
    Image8 g(disp.width(), disp.height(), 1);
    for (int y = 0; y < disp.height(); ++y)
        for (int x = 0; x < disp.width(); ++x) {
            const float d = disp.at(x, y);
            g.at(x, y) = d < 0.f ? 0 : static_cast<uint8_t>(clampi(static_cast<int>(d * k + 0.5f), 0, 255));
        }
    return save_pgm(path, g);
}

}  // namespace apg

// Texture files for material textures: TGA (uncompressed or RLE; 8-bit grey,
// 24- and 32-bit colour), binary PPM (P6, maxval 255) and PFM (PF colour, Pf
// grey).  8-bit images become GPUGI_TEXTURE_RGBA8_SRGB texels, PFM images
// GPUGI_TEXTURE_RGBA32F, alpha 1 where the file has none.  Rows are stored
// bottom row first, the order an OpenGL texture upload of a bottom-up image
// takes: TGA's origin bit and the top-down PPM rows are flipped to it, PFM is
// bottom-up already.
#include <cctype>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iterator>
#include <string>
#include <vector>

#include "scene.hpp"

namespace gpugi {

namespace {

bool EndsWith(const std::string& s, const char* suffix) {
  const size_t n = std::strlen(suffix);
  if (s.size() < n) return false;
  for (size_t i = 0; i < n; ++i)
    if (std::tolower((unsigned char)s[s.size() - n + i]) != suffix[i]) return false;
  return true;
}

std::string LoadTga(const std::vector<uint8_t>& f, HostTexture* out) {
  if (f.size() < 18) return "truncated TGA header";
  const int idLen = f[0], cmapType = f[1], type = f[2];
  const int width = f[12] | (f[13] << 8), height = f[14] | (f[15] << 8), bpp = f[16], desc = f[17];
  if (cmapType != 0) return "colour-mapped TGA is not supported";
  const bool rle = type == 10 || type == 11, grey = type == 3 || type == 11;
  if (!(type == 2 || type == 3 || type == 10 || type == 11)) return "unsupported TGA image type " + std::to_string(type);
  if (width <= 0 || height <= 0) return "empty TGA image";
  if (grey ? bpp != 8 : (bpp != 24 && bpp != 32)) return "unsupported TGA pixel depth " + std::to_string(bpp);
  const int bytes = bpp / 8;
  const size_t n = (size_t)width * height;
  std::vector<uint8_t> px(n * bytes);
  size_t pos = 18 + (size_t)idLen;
  if (!rle) {
    if (f.size() < pos + px.size()) return "truncated TGA pixel data";
    std::memcpy(px.data(), f.data() + pos, px.size());
  } else {
    size_t i = 0;
    while (i < n) {
      if (pos >= f.size()) return "truncated TGA RLE data";
      const int head = f[pos++];
      const size_t count = (size_t)(head & 0x7f) + 1;
      if (i + count > n) return "TGA RLE packet past the image";
      if (head & 0x80) {
        if (pos + bytes > f.size()) return "truncated TGA RLE data";
        for (size_t k = 0; k < count; ++k) std::memcpy(&px[(i + k) * bytes], &f[pos], bytes);
        pos += bytes;
      } else {
        if (pos + count * bytes > f.size()) return "truncated TGA RLE data";
        std::memcpy(&px[i * bytes], &f[pos], count * bytes);
        pos += count * bytes;
      }
      i += count;
    }
  }
  const bool topDown = (desc & 0x20) != 0, rightLeft = (desc & 0x10) != 0;
  out->width = width;
  out->height = height;
  out->format = GPUGI_TEXTURE_RGBA8_SRGB;
  out->data.assign(n * 4, 255);
  for (int y = 0; y < height; ++y)
    for (int x = 0; x < width; ++x) {
      const int sy = topDown ? height - 1 - y : y, sx = rightLeft ? width - 1 - x : x;
      const uint8_t* s = &px[((size_t)sy * width + sx) * bytes];
      uint8_t* d = &out->data[((size_t)y * width + x) * 4];
      if (grey) {
        d[0] = d[1] = d[2] = s[0];
      } else {  // stored BGR(A)
        d[0] = s[2];
        d[1] = s[1];
        d[2] = s[0];
        if (bytes == 4) d[3] = s[3];
      }
    }
  return std::string();
}

// Whitespace-separated header tokens ('#' comments allowed), then exactly one
// whitespace character before the raster.
bool HeaderTokens(const std::vector<uint8_t>& f, int count, std::vector<std::string>* tok, size_t* pos) {
  size_t p = 2;
  while ((int)tok->size() < count) {
    while (p < f.size() && (std::isspace(f[p]) || f[p] == '#')) {
      if (f[p] == '#')
        while (p < f.size() && f[p] != '\n') ++p;
      else
        ++p;
    }
    std::string t;
    while (p < f.size() && !std::isspace(f[p])) t.push_back((char)f[p++]);
    if (t.empty()) return false;
    tok->push_back(t);
  }
  if (p >= f.size()) return false;
  *pos = p + 1;
  return true;
}

std::string LoadPpm(const std::vector<uint8_t>& f, HostTexture* out) {
  std::vector<std::string> tok;
  size_t pos = 0;
  if (!HeaderTokens(f, 3, &tok, &pos)) return "bad PPM header";
  const int width = std::atoi(tok[0].c_str()), height = std::atoi(tok[1].c_str());
  if (width <= 0 || height <= 0 || std::atoi(tok[2].c_str()) != 255) return "PPM must be P6 with maxval 255";
  const size_t n = (size_t)width * height;
  if (f.size() < pos + n * 3) return "truncated PPM raster";
  out->width = width;
  out->height = height;
  out->format = GPUGI_TEXTURE_RGBA8_SRGB;
  out->data.assign(n * 4, 255);
  for (int y = 0; y < height; ++y)
    for (int x = 0; x < width; ++x) {
      const uint8_t* s = &f[pos + ((size_t)(height - 1 - y) * width + x) * 3];  // PPM rows run top-down
      uint8_t* d = &out->data[((size_t)y * width + x) * 4];
      d[0] = s[0];
      d[1] = s[1];
      d[2] = s[2];
    }
  return std::string();
}

std::string LoadPfmTexture(const std::vector<uint8_t>& f, HostTexture* out) {
  const bool colour = f[1] == 'F';
  std::vector<std::string> tok;
  size_t pos = 0;
  if (!HeaderTokens(f, 3, &tok, &pos)) return "bad PFM header";
  const int width = std::atoi(tok[0].c_str()), height = std::atoi(tok[1].c_str());
  const double scale = std::atof(tok[2].c_str());
  if (width <= 0 || height <= 0 || scale == 0.0) return "bad PFM header";
  const int ch = colour ? 3 : 1;
  const size_t n = (size_t)width * height;
  if (f.size() < pos + n * ch * 4) return "truncated PFM raster";
  out->width = width;
  out->height = height;
  out->format = GPUGI_TEXTURE_RGBA32F;
  out->data.assign(n * 16, 0);
  float* d = reinterpret_cast<float*>(out->data.data());
  for (size_t i = 0; i < n; ++i) {
    float v[3];
    for (int c = 0; c < ch; ++c) {
      uint32_t u;
      std::memcpy(&u, &f[pos + (i * ch + c) * 4], 4);
      if (scale > 0.0) u = (u >> 24) | ((u >> 8) & 0xff00u) | ((u << 8) & 0xff0000u) | (u << 24);  // big-endian
      std::memcpy(&v[c], &u, 4);
    }
    d[4 * i + 0] = v[0];
    d[4 * i + 1] = colour ? v[1] : v[0];
    d[4 * i + 2] = colour ? v[2] : v[0];
    d[4 * i + 3] = 1.0f;
  }
  return std::string();
}

}  // namespace

std::string LoadTextureFile(const std::string& path, HostTexture* out) {
  std::ifstream in(path, std::ios::binary);
  if (!in) return "cannot open texture " + path;
  const std::vector<uint8_t> f((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  std::string err;
  if (f.size() >= 2 && f[0] == 'P' && f[1] == '6')
    err = LoadPpm(f, out);
  else if (f.size() >= 2 && f[0] == 'P' && (f[1] == 'F' || f[1] == 'f'))
    err = LoadPfmTexture(f, out);
  else if (EndsWith(path, ".tga"))
    err = LoadTga(f, out);
  else
    err = "unsupported texture format (TGA, binary PPM or PFM)";
  if (!err.empty()) return path + ": " + err;
  out->file = path;
  return std::string();
}

}  // namespace gpugi

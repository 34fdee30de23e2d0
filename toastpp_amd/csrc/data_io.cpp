// Data file formats of the reference around the forward model: nodal images
// (NIM) and the bracketed vector text of the .fem data files.  Host-only.
//
//   NIM:   "NIM" (or "RIM"), header lines "Key = value" up to "EndHeader"
//          (ImageSize gives the image length), then blocks "Image <k>"
//          followed by ImageSize whitespace-separated values
//          (fwdfem.cc:297-347, toastmodule.cc:905-973).
//   .fem:  one vector "[v0 v1 ...]" (TVector operator<< / >>,
//          vector_imp.hpp:1359-1420; WriteData, fwdfem.cc:457-462).
#include <cctype>
#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "common.hpp"

namespace tmi {

static std::string trim(const std::string &s)
{
    size_t a = 0, b = s.size();
    while (a < b && isspace((unsigned char)s[a])) a++;
    while (b > a && isspace((unsigned char)s[b - 1])) b--;
    return s.substr(a, b - a);
}

static bool starts_nocase(const std::string &s, const char *p)
{
    size_t n = strlen(p);
    return s.size() >= n && strncasecmp(s.c_str(), p, n) == 0;
}

// Parse one number token like an istream >> double would, rejecting trailing junk.
static bool parse_double(const std::string &tok, double &v)
{
    if (tok.empty()) return false;
    char *end = nullptr;
    errno = 0;
    v = strtod(tok.c_str(), &end);
    return end && *end == '\0';
}

struct Nim {
    std::string mesh;
    int size = 0;
    std::vector<double> image;   // the selected image
};

// Read a NIM file and keep image `want` (0-based); want < 0 or beyond the
// last image selects the last one, as toast_read_nim (toastmodule.cc:922-931)
// overwrites its buffer with every image until the index matches or the
// file ends.
static Nim read_nim(const char *path, int &nimg_total, int want)
{
    std::ifstream f(path);
    TMI_CHECK(f.good(), TMI_ERR_IO, "cannot open NIM file '%s'", path);
    std::string line;
    TMI_CHECK((bool)std::getline(f, line), TMI_ERR_IO, "'%s' is empty", path);
    line = trim(line);
    TMI_CHECK(line == "NIM" || line == "RIM", TMI_ERR_IO, "'%s' is not a NIM file (first line '%s')", path,
              line.c_str());
    Nim nim;
    bool end = false;
    while (std::getline(f, line)) {
        line = trim(line);
        if (strcasecmp(line.c_str(), "EndHeader") == 0) { end = true; break; }
        size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        std::string key = trim(line.substr(0, eq)), val = trim(line.substr(eq + 1));
        if (strcasecmp(key.c_str(), "ImageSize") == 0) nim.size = atoi(val.c_str());
        else if (strcasecmp(key.c_str(), "Mesh") == 0) nim.mesh = val;
    }
    TMI_CHECK(end, TMI_ERR_IO, "'%s': header has no EndHeader line", path);
    TMI_CHECK(nim.size > 0, TMI_ERR_IO, "'%s': header gives no positive ImageSize", path);
    nimg_total = 0;
    std::string tok;
    for (;;) {
        // skip to the next "Image" line
        bool found = false;
        while (std::getline(f, line)) {
            if (starts_nocase(trim(line), "Image")) { found = true; break; }
        }
        if (!found) break;
        bool store = want < 0 || nimg_total <= want;
        std::vector<double> scratch;
        std::vector<double> &img = store ? nim.image : scratch;
        img.resize((size_t)nim.size);
        for (int i = 0; i < nim.size; i++) {
            TMI_CHECK((bool)(f >> tok), TMI_ERR_IO, "'%s': image %d ends after %d of %d values", path, nimg_total, i,
                      nim.size);
            TMI_CHECK(parse_double(tok, img[i]), TMI_ERR_IO, "'%s': image %d value %d ('%s') is not a number", path,
                      nimg_total, i, tok.c_str());
        }
        std::getline(f, line);   // rest of the value line
        nimg_total++;
    }
    TMI_CHECK(nimg_total > 0, TMI_ERR_IO, "'%s' holds no image", path);
    return nim;
}

static void read_vector(const char *path, std::vector<double> &v)
{
    std::ifstream f(path);
    TMI_CHECK(f.good(), TMI_ERR_IO, "cannot open vector file '%s'", path);
    std::stringstream ss;
    ss << f.rdbuf();
    std::string s = ss.str();
    size_t a = s.find('[');
    TMI_CHECK(a != std::string::npos, TMI_ERR_IO, "'%s': no '[' opening the vector", path);
    size_t b = s.find(']', a);
    TMI_CHECK(b != std::string::npos, TMI_ERR_IO, "'%s': no ']' closing the vector", path);
    std::istringstream in(s.substr(a + 1, b - a - 1));
    std::string tok;
    v.clear();
    while (in >> tok) {
        double x;
        TMI_CHECK(parse_double(tok, x), TMI_ERR_IO, "'%s': entry %zu ('%s') is not a number", path, v.size(),
                  tok.c_str());
        v.push_back(x);
    }
}

}  // namespace tmi

using namespace tmi;

#define API_BEGIN try {
#define API_END                         \
    return TMI_OK;                      \
    }                                   \
    catch (const tmi::Error &e)         \
    {                                   \
        set_last_error(e.what());       \
        return e.code;                  \
    }                                   \
    catch (const std::exception &e)     \
    {                                   \
        set_last_error(e.what());       \
        return TMI_ERR_IO;              \
    }

extern "C" {

int tmi_read_nim(const char *path, int idx, int *size, int *nimg, double *img)
{
    API_BEGIN
    TMI_CHECK(path != nullptr, TMI_ERR_INVALID_ARG, "path must not be NULL");
    int total = 0;
    Nim nim = read_nim(path, total, idx);
    if (size) *size = nim.size;
    if (nimg) *nimg = total;
    if (img) memcpy(img, nim.image.data(), sizeof(double) * nim.size);
    API_END
}

int tmi_write_nim(const char *path, const char *meshname, int size, int nimg, const double *imgs, int style)
{
    API_BEGIN
    TMI_CHECK(path && imgs, TMI_ERR_INVALID_ARG, "path and imgs must not be NULL");
    TMI_CHECK(size > 0 && nimg > 0, TMI_ERR_INVALID_ARG, "image size and count must be positive");
    TMI_CHECK(style == TMI_NIM_FWDFEM || style == TMI_NIM_TOASTMOD, TMI_ERR_INVALID_ARG, "unknown NIM style %d",
              style);
    FILE *f = fopen(path, "w");
    TMI_CHECK(f != nullptr, TMI_ERR_IO, "cannot create '%s'", path);
    fprintf(f, "NIM\nMesh = %s\nSolutionType = N/A\nImageSize = %d\nEndHeader\n", meshname ? meshname : "", size);
    const char *fmt = style == TMI_NIM_FWDFEM ? "%.10g " : "%.12e ";
    for (int k = 0; k < nimg; k++) {
        fprintf(f, "Image %d\n", k);
        const double *v = imgs + (size_t)k * size;
        for (int i = 0; i < size; i++) fprintf(f, fmt, v[i]);
        fputc('\n', f);
    }
    bool ok = !ferror(f);
    ok = (fclose(f) == 0) && ok;
    TMI_CHECK(ok, TMI_ERR_IO, "write to '%s' failed", path);
    API_END
}

int tmi_read_vector(const char *path, int *n, double *v)
{
    API_BEGIN
    TMI_CHECK(path != nullptr, TMI_ERR_INVALID_ARG, "path must not be NULL");
    std::vector<double> x;
    read_vector(path, x);
    if (v) {
        TMI_CHECK(n != nullptr && *n >= (int)x.size(), TMI_ERR_INVALID_ARG,
                  "'%s' holds %zu values, more than the buffer's %d", path, x.size(), n ? *n : 0);
        memcpy(v, x.data(), sizeof(double) * x.size());
    }
    if (n) *n = (int)x.size();
    API_END
}

int tmi_write_vector(const char *path, int n, const double *v)
{
    API_BEGIN
    TMI_CHECK(path && (v || n == 0) && n >= 0, TMI_ERR_INVALID_ARG, "invalid vector");
    FILE *f = fopen(path, "w");
    TMI_CHECK(f != nullptr, TMI_ERR_IO, "cannot create '%s'", path);
    fputc('[', f);
    for (int i = 0; i < n; i++) fprintf(f, i ? " %.10e" : "%.10e", v[i]);
    fputs("]\n", f);
    bool ok = !ferror(f);
    ok = (fclose(f) == 0) && ok;
    TMI_CHECK(ok, TMI_ERR_IO, "write to '%s' failed", path);
    API_END
}

}  // extern "C"

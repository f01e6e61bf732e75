// vct_txt.h -- reader for the reference's positional input-txt files.
//
// Format (M/projectionDataTOMO.cpp:17-64, F/projectionDataTOMO.cpp:17-51):
// six header lines, then for every field two skipped lines (blank line and
// comment) followed by the value line, parsed with sscanf("%d" / "%lf" /
// "%s" / "%lf %lf %lf").  A missing line leaves the field at its previous
// value, as fgets/sscanf do there.
#pragma once

#include <cstdio>
#include <cstring>
#include <string>

namespace vct {

class TxtReader {
public:
    explicit TxtReader(const char* path) { fp_ = std::fopen(path, "r"); }
    ~TxtReader()
    {
        if (fp_) std::fclose(fp_);
    }
    bool ok() const { return fp_ != nullptr; }
    void skip(int n)
    {
        for (int i = 0; i < n; i++) line();
    }
    void read_int(int* v)
    {
        skip(2);
        std::sscanf(line(), "%d", v);
    }
    void read_double(double* v)
    {
        skip(2);
        std::sscanf(line(), "%lf", v);
    }
    void read_double3(double* a, double* b, double* c)
    {
        skip(2);
        std::sscanf(line(), "%lf %lf %lf", a, b, c);
    }
    void read_string(std::string* s)
    {
        skip(2);
        char tmp[512] = {0};
        if (std::sscanf(line(), "%511s", tmp) == 1) *s = tmp;
    }

private:
    // At end of file fgets leaves the buffer untouched, so the previous line
    // is parsed again (the reference's static lineBuf behaves the same way).
    const char* line()
    {
        if (fp_) (void)std::fgets(buf_, sizeof(buf_), fp_);
        return buf_;
    }
    FILE* fp_ = nullptr;
    char buf_[512] = {0};
};

}  // namespace vct

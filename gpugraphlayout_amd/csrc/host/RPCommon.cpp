#include "RPCommon.hpp"
#include <cmath>
#include <cstdlib>
#include <sys/stat.h>

bool is_file_exists(std::string filepath)
{
    struct stat st;
    return stat(filepath.c_str(), &st) == 0;
}

std::string basename(std::string filepath)
{
    while (filepath.size() > 1 && filepath.back() == '/') filepath.pop_back();
    size_t k = filepath.find_last_of('/');
    return k == std::string::npos ? filepath : filepath.substr(k + 1);
}

namespace RPGraph
{
    float get_random(float lowerbound, float upperbound)
    {
        const float r = static_cast<float>(random());
        return lowerbound + (upperbound - lowerbound) * r / static_cast<float>(RAND_MAX);
    }

    float Real2DVector::magnitude() const { return std::sqrt(x * x + y * y); }
}

// h5_dump.cpp — reads items of an output.h5 with the HDF5 C++ API (the API
// the reference's report.cpp writes with) and prints one line per item, for
// tests/test_report_h5.py.  Items: u:<dataset> (unsigned long long scalar),
// d:<dataset> (double scalar), s:<dataset> (string scalar), v:<dataset>
// (double array: its dims, then its values), g:<group>:<attribute> and
// t:<dataset>:<attribute> (string attributes).
#include <H5Cpp.h>

#include <cstdio>
#include <string>
#include <vector>

static std::string attr_string(const H5::H5Object& o, const std::string& name)
{
    H5::Attribute a = o.openAttribute(name);
    std::string v;
    a.read(a.getStrType(), v);
    return v;
}

int main(int argc, char** argv)
{
    if (argc < 2) return 2;
    try {
        H5::H5File f(argv[1], H5F_ACC_RDONLY);
        for (int i = 2; i < argc; ++i) {
            const std::string item(argv[i]);
            const char kind = item[0];
            const std::string path = item.substr(2);
            if (kind == 'u') {
                unsigned long long v = 0;
                f.openDataSet(path).read(&v, H5::PredType::NATIVE_ULLONG);
                std::printf("%s %llu\n", item.c_str(), v);
            } else if (kind == 'd') {
                double v = 0;
                f.openDataSet(path).read(&v, H5::PredType::NATIVE_DOUBLE);
                std::printf("%s %.17g\n", item.c_str(), v);
            } else if (kind == 's') {
                std::string v;
                H5::DataSet d = f.openDataSet(path);
                d.read(v, d.getStrType());
                std::printf("%s %s\n", item.c_str(), v.c_str());
            } else if (kind == 'v') {
                H5::DataSet d = f.openDataSet(path);
                H5::DataSpace sp = d.getSpace();
                const int rank = sp.getSimpleExtentNdims();
                std::vector<hsize_t> dims(rank);
                sp.getSimpleExtentDims(dims.data());
                size_t n = 1;
                for (hsize_t x : dims) n *= (size_t)x;
                std::vector<double> v(n);
                d.read(v.data(), H5::PredType::NATIVE_DOUBLE);
                std::printf("%s dims", item.c_str());
                for (hsize_t x : dims) std::printf(" %llu", (unsigned long long)x);
                std::printf(" values");
                for (double x : v) std::printf(" %.17g", x);
                std::printf("\n");
            } else if (kind == 'g' || kind == 't') {
                const size_t c = path.rfind(':');
                const std::string obj = path.substr(0, c), att = path.substr(c + 1);
                const std::string v = kind == 'g' ? attr_string(f.openGroup(obj), att)
                                                  : attr_string(f.openDataSet(obj), att);
                std::printf("%s %s\n", item.c_str(), v.c_str());
            }
        }
    } catch (const H5::Exception& e) {
        std::fprintf(stderr, "HDF5: %s\n", e.getCDetailMsg());
        return 1;
    }
    return 0;
}

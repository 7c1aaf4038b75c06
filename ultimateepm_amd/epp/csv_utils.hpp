// csv_utils.hpp — the column CSV writer the reference's DOS tool uses
// (uepm::utils::export_multiple_vector_to_csv, src/Common/csv_utils.hpp:21-50):
// header "name_0,name_1,...,DumbColumn", one row per index with the columns'
// values in default stream precision and a trailing 0.  Columns of unequal
// length write nothing.
#pragma once

#include <fstream>
#include <iostream>
#include <string>
#include <vector>

namespace uepm::utils {

inline void export_multiple_vector_to_csv(const std::string& filename, const std::vector<std::string>& header_columns,
                                          const std::vector<std::vector<double>>& columns) {
    if (columns.empty()) return;
    const std::size_t rows = columns[0].size();
    for (const auto& c : columns)
        if (c.size() != rows) {
            std::cout << "ERROR WHEN EXPORTING VECTORS TO : " << filename << ", mismatch between vector sizes : " << rows
                      << " != " << c.size() << std::endl;
            return;
        }
    std::ofstream out(filename);
    for (const auto& name : header_columns) out << name << ",";
    out << "DumbColumn\n";
    for (std::size_t r = 0; r < rows; ++r) {
        for (const auto& c : columns) out << c[r] << ",";
        out << 0.0 << "\n";
    }
}

}  // namespace uepm::utils

// Timings and checksums of the measured rounds of one experiment, printed as the
// reference's CSV line (include/experiments/experiment_report.hpp:82-93: the run
// parameters, then average_time_ms, average_time_per_cell_ps, std_time_ms,
// rounds_had_same_checksums, checksum) and as a readable summary.
#ifndef CELLATO_EXPERIMENTS_EXPERIMENT_REPORT_HPP
#define CELLATO_EXPERIMENTS_EXPERIMENT_REPORT_HPP

#include <cmath>
#include <ostream>
#include <string>
#include <vector>

#include "run_params.hpp"

namespace cellato::run {

struct experiment_report {
    run_params params;
    std::vector<double> execution_times_ms;   // one per measured round
    std::vector<std::string> checksums;

    static std::string csv_header() {
        return run_params::csv_header() + ",average_time_ms,average_time_per_cell_ps,std_time_ms,rounds_had_same_checksums,checksum";
    }

    std::string csv_line() const {
        return params.csv_line() + "," + std::to_string(average_time_ms()) + "," +
               std::to_string(average_time_per_cell_ps()) + "," + std::to_string(std_time_ms()) + "," +
               (rounds_had_same_checksums() ? "true" : "false") + "," + (checksums.empty() ? "" : checksums.front());
    }

    double average_time_ms() const {
        if (execution_times_ms.empty()) return 0.0;
        double s = 0.0;
        for (double t : execution_times_ms) s += t;
        return s / static_cast<double>(execution_times_ms.size());
    }

    // population standard deviation over the measured rounds
    double std_time_ms() const {
        if (execution_times_ms.empty()) return 0.0;
        const double m = average_time_ms();
        double s = 0.0;
        for (double t : execution_times_ms) s += (t - m) * (t - m);
        return std::sqrt(s / static_cast<double>(execution_times_ms.size()));
    }

    bool rounds_had_same_checksums() const {
        for (const auto& c : checksums)
            if (c != checksums.front()) return false;
        return true;
    }

    // picoseconds per cell update
    double average_time_per_cell_ps() const {
        if (params.x_size == 0 || params.y_size == 0 || params.steps == 0) return 0.0;
        return average_time_ms() * 1e9 / (static_cast<double>(params.x_size) * params.y_size) / params.steps;
    }

    void pretty_print(std::ostream& os) const {
        const double cells = static_cast<double>(params.x_size) * params.y_size;
        os << "==============================================\n"
           << "          EXPERIMENT REPORT\n"
           << "==============================================\n"
           << "Grid Size: " << params.x_size << " x " << params.y_size << " (" << static_cast<long long>(cells)
           << " cells), steps " << params.steps << ", seed " << params.seed << "\n"
           << "Avg. Execution Time: " << average_time_ms() << " ms (std " << std_time_ms() << " ms, "
           << execution_times_ms.size() << " rounds)\n"
           << "Avg. Time Per Cell: " << average_time_per_cell_ps() << " ps";
        if (average_time_ms() > 0) os << " (" << cells * params.steps / (average_time_ms() * 1e-3) / 1e9 << " GCUPS)";
        os << "\n" << (rounds_had_same_checksums() ? "All rounds produced identical results.\n"
                                                   : "Warning: different rounds produced different results!\n");
        if (!checksums.empty()) os << "Checksum: " << checksums.front() << "\n";
        os << "==============================================\n";
    }
};

} // namespace cellato::run

#endif

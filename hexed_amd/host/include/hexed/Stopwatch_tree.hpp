// Wall-clock timing tree: the interface Kernel and Solver use from Stopwatch.hpp (Stopwatch.hpp:14-45) and
// Stopwatch_tree.hpp (Stopwatch_tree.hpp:15-27) in the hexed tree. Device work is asynchronous, so the timed kernel
// overloads synchronize the device before their stopwatches stop.
#ifndef HEXED_AMD_STOPWATCH_TREE_HPP_
#define HEXED_AMD_STOPWATCH_TREE_HPP_

#include <chrono>
#include <cstdio>
#include <map>
#include <stdexcept>
#include <string>

namespace hexed
{

class Stopwatch
{
  std::chrono::steady_clock::time_point t0;
  double total = 0.;
  bool is_running = false;
  int calls = 0;

  public:
  //! starts the stopwatch on construction and pauses it on destruction
  class Operator
  {
    Stopwatch& sw;

    public:
    Operator(Stopwatch& s) : sw{s} { sw.start(); }
    ~Operator() { sw.pause(); }
  };
  //! throws if already running
  void start()
  {
    if (is_running) throw std::runtime_error("attempt to start stopwatch which is already running");
    t0 = std::chrono::steady_clock::now();
    is_running = true;
  }
  //! stops and adds the elapsed time; a stopwatch that is not running is left alone (this runs in destructors)
  void pause()
  {
    if (!is_running) return;
    total += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    is_running = false;
    ++calls;
  }
  void reset()
  {
    total = 0.;
    calls = 0;
  }
  bool running() const { return is_running; }
  int n_calls() const { return calls; }
  double time() const { return total; }
  Stopwatch operator+(Stopwatch other) const
  {
    if (is_running || other.is_running) throw std::runtime_error("cannot add running stopwatches");
    other.total += total;
    other.calls += calls;
    return other;
  }
  Stopwatch& operator+=(Stopwatch other) { return *this = *this + other; }
};

class Stopwatch_tree
{
  std::string report_at(const std::string& indent) const
  {
    std::string out;
    if (work_unit_name.empty()) {
      out = "Total:\n";
    } else {
      char line[512];
      std::snprintf(line, sizeof(line), "%i %ss completed in %g s", work_units_completed, work_unit_name.c_str(),
                    stopwatch.time());
      out = line;
      if (work_units_completed) {
        std::snprintf(line, sizeof(line), " at %g s / %s", stopwatch.time() / work_units_completed,
                      work_unit_name.c_str());
        out += line;
      }
      out += ".\n";
    }
    for (const auto& [name, child] : children) out += indent + "    " + name + ": " + child.report_at(indent + "    ");
    return out;
  }

  public:
  Stopwatch stopwatch;
  std::map<std::string, Stopwatch_tree> children;
  int work_units_completed = 0;
  std::string work_unit_name;
  Stopwatch_tree(std::string work_unit_name_arg = "", std::map<std::string, Stopwatch_tree> init_children = {})
  : children{std::move(init_children)}, work_unit_name{std::move(work_unit_name_arg)}
  {}
  //! human-readable summary: units completed, total time and time per unit, children indented below
  std::string report() const { return report_at(""); }
};

}  // namespace hexed
#endif

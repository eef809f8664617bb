// Wall-clock timing tree (the interface Kernel and Solver use from Stopwatch.hpp / Stopwatch_tree.hpp:15-27 in the
// hexed tree). Device work is asynchronous, so the timed kernel overloads synchronize the device before stopping.
#ifndef HEXED_AMD_STOPWATCH_TREE_HPP_
#define HEXED_AMD_STOPWATCH_TREE_HPP_

#include <chrono>
#include <map>
#include <string>

namespace hexed
{

class Stopwatch
{
  std::chrono::steady_clock::time_point t0;
  double total = 0.;
  bool running = false;
  int n_calls = 0;

  public:
  void start()
  {
    t0 = std::chrono::steady_clock::now();
    running = true;
  }
  void pause()
  {
    if (!running) return;
    total += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    running = false;
    ++n_calls;
  }
  double time() const { return total; }
  int n_calls_made() const { return n_calls; }
  //! starts the stopwatch on construction and pauses it on destruction
  class Operator
  {
    Stopwatch& sw;
    public:
    explicit Operator(Stopwatch& s) : sw{s} { sw.start(); }
    ~Operator() { sw.pause(); }
  };
};

class Stopwatch_tree
{
  public:
  std::string work_unit_name;
  Stopwatch stopwatch;
  long long work_units_completed = 0;
  std::map<std::string, Stopwatch_tree> children;
  explicit Stopwatch_tree(std::string unit = "") : work_unit_name{std::move(unit)} {}
};

}  // namespace hexed
#endif

// pybind11 module vecsearch_amd._vecsearch: the reference's GPU binding surface
// (python/bindings.cpp:18-56, 151-197) over the MI355X engine.
//
// Kept from the reference: FloatArray forcecast (float64 -> float32 copy, :18), 1-D queries
// become (1, d) (:174), check_2d -> ValueError (:44-47), mask parsing as_mask (:28-36),
// parse_metric (:38-42), ids as int64 with -1 for missing (:49-56), the GIL released around
// add/search (:161, :179-182), method names auto/naive/skinny/tiled (:168-173).
// Added: search(queries, k, method, filter=None) with the HNSW binding's mask convention
// (:84-91), where a mask given third is the filter, and search_batch for per-query k and
// per-query allowed-id lists.
#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>

#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include "vecsearch/gpu_brute_force.h"
#include "vsg.h"

namespace py = pybind11;
using namespace vecsearch;
using FloatArray = py::array_t<float, py::array::c_style | py::array::forcecast>;
using MaskArray = py::array_t<uint8_t, py::array::c_style | py::array::forcecast>;

namespace {

struct Mask {
  MaskArray arr;
  const uint8_t* ptr = nullptr;
};

Mask as_mask(const py::object& filter, size_t n) {
  Mask m;
  if (filter.is_none()) return m;
  m.arr = MaskArray::ensure(filter);
  if (!m.arr || m.arr.ndim() != 1 || size_t(m.arr.shape(0)) != n)
    throw std::invalid_argument("filter must be a 1-D bool array of length " + std::to_string(n));
  m.ptr = m.arr.data();
  return m;
}

Metric parse_metric(const std::string& s) {
  if (s == "l2") return Metric::L2;
  if (s == "ip" || s == "cosine") return Metric::InnerProduct;
  throw std::invalid_argument(
      "metric must be 'l2', 'ip' or 'cosine' (cosine expects normalized vectors)");
}

GpuMethod parse_method(const std::string& method) {
  if (method == "auto") return GpuMethod::Auto;
  if (method == "naive") return GpuMethod::Naive;
  if (method == "skinny") return GpuMethod::Skinny;
  if (method == "tiled") return GpuMethod::Tiled;
  if (method == "bf16") return GpuMethod::Bf16;
  if (method == "fp8") return GpuMethod::Fp8;
  throw std::invalid_argument("method must be 'auto', 'naive', 'skinny', 'tiled', 'bf16' or 'fp8'");
}

void check_2d(const FloatArray& a, size_t dim) {
  if (a.ndim() != 2 || size_t(a.shape(1)) != dim)
    throw std::invalid_argument("expected array of shape (n, " + std::to_string(dim) + ")");
}

py::array_t<int64_t> ids_to_int64(const py::array_t<uint32_t>& ids) {
  py::array_t<int64_t> out({ids.shape(0), ids.shape(1)});
  const uint32_t* src = ids.data();
  int64_t* dst = out.mutable_data();
  for (py::ssize_t i = 0; i < ids.size(); ++i)
    dst[i] = src[i] == 0xFFFFFFFFu ? -1 : int64_t(src[i]);
  return out;
}

// A captured search (vsg_graph_*): raw device pointers, replayed by launch().
class PyGraph {
 public:
  PyGraph(py::object index, uintptr_t d_q, size_t nq, size_t k, uintptr_t d_ids, uintptr_t d_dists,
          const std::string& method)
      : index_(std::move(index)) {
    auto& idx = index_.cast<GpuBruteForceIndex&>();
    if (vsg_graph_create(idx.handle(), reinterpret_cast<const float*>(d_q), nq, k,
                         static_cast<int>(parse_method(method)), reinterpret_cast<uint32_t*>(d_ids),
                         reinterpret_cast<float*>(d_dists), &g_) != VSG_OK)
      throw std::invalid_argument(vsg_last_error());
  }
  // A captured filtered batch (vsg_graph_create_lists); ks / ranges / allow_ids as search_batch.
  PyGraph(py::object index, uintptr_t d_q, py::array_t<uint32_t, py::array::c_style | py::array::forcecast> ks,
          py::array_t<uint64_t, py::array::c_style | py::array::forcecast> ranges,
          py::array_t<uint32_t, py::array::c_style | py::array::forcecast> allow_ids, size_t kmax, uintptr_t d_ids,
          uintptr_t d_dists)
      : index_(std::move(index)) {
    auto& idx = index_.cast<GpuBruteForceIndex&>();
    if (ks.ndim() != 1) throw std::invalid_argument("ks must be a 1-D array with one k per query");
    const size_t nq = size_t(ks.shape(0));
    if (ranges.ndim() != 2 || size_t(ranges.shape(0)) != nq || ranges.shape(1) != 2)
      throw std::invalid_argument("ranges must have shape (nq, 2): [begin, end) into allow_ids");
    if (allow_ids.ndim() != 1) throw std::invalid_argument("allow_ids must be a 1-D array");
    for (size_t i = 0; i < nq; ++i) kmax = std::max<size_t>(kmax, ks.data()[i]);
    if (vsg_graph_create_lists(idx.handle(), reinterpret_cast<const float*>(d_q), nq, ks.data(), kmax,
                               ranges.data(), allow_ids.data(), size_t(allow_ids.shape(0)),
                               reinterpret_cast<uint32_t*>(d_ids), reinterpret_cast<float*>(d_dists), &g_) != VSG_OK)
      throw std::invalid_argument(vsg_last_error());
  }
  ~PyGraph() { close(); }
  void launch(uintptr_t stream) {
    if (!g_) throw std::runtime_error("graph is closed");
    if (vsg_graph_launch(g_, reinterpret_cast<void*>(stream)) != VSG_OK) throw std::runtime_error(vsg_last_error());
  }
  void close() {
    if (g_) vsg_graph_destroy(g_);
    g_ = nullptr;
  }

 private:
  py::object index_;
  vsg_graph* g_ = nullptr;
};

// Python face of the serving loop (vsg_server_*): submit() keeps the request's arrays alive
// until wait() returns them.
class PyServer {
 public:
  PyServer(py::object index, size_t max_batch_queries, size_t workspace_bytes, int slots,
           const std::string& method, bool tune)
      : index_(std::move(index)) {
    auto& idx = index_.cast<GpuBruteForceIndex&>();
    dim_ = idx.dim();
    vsg_server_config cfg{};
    cfg.max_batch_queries = max_batch_queries;
    cfg.workspace_bytes = workspace_bytes;
    cfg.slots = slots;
    cfg.method = static_cast<int>(parse_method(method));
    cfg.tune_mode = tune ? 0 : 1;
    if (vsg_server_create(idx.handle(), &cfg, &s_) != VSG_OK) throw std::invalid_argument(vsg_last_error());
  }
  ~PyServer() { close(); }

  uint64_t submit(FloatArray q, size_t k, const py::object& allow) {
    if (!s_) throw std::runtime_error("server is closed");
    if (q.ndim() == 1) q = q.reshape({py::ssize_t(1), q.shape(0)});
    check_2d(q, dim_);
    Pending p;
    p.q = q;
    const size_t nq = q.shape(0);
    const uint32_t* ap = nullptr;
    size_t na = 0;
    if (!allow.is_none()) {
      p.allow = py::array_t<uint32_t, py::array::c_style | py::array::forcecast>::ensure(allow);
      if (!p.allow || p.allow.ndim() != 1) throw std::invalid_argument("allow_ids must be a 1-D array of ids");
      ap = p.allow.data();
      na = size_t(p.allow.shape(0));
      static const uint32_t kEmpty = 0;
      if (na == 0) ap = &kEmpty;  // an empty allowed set is still a filter
    }
    p.ids = py::array_t<uint32_t>({nq, k});
    p.dists = py::array_t<float>({nq, k});
    uint64_t t = 0;
    if (vsg_server_submit(s_, p.q.data(), nq, k, ap, na, p.ids.mutable_data(), p.dists.mutable_data(), &t) != VSG_OK)
      throw std::invalid_argument(vsg_last_error());
    pending_.emplace(t, std::move(p));
    return t;
  }

  py::tuple wait(uint64_t ticket) {
    auto it = pending_.find(ticket);
    if (it == pending_.end()) throw std::invalid_argument("unknown or already waited ticket");
    Pending p = std::move(it->second);
    pending_.erase(it);
    double ms = 0;
    int rc;
    {
      py::gil_scoped_release nogil;
      rc = vsg_server_wait(s_, ticket, &ms);
    }
    if (rc != VSG_OK) throw std::runtime_error(vsg_last_error());
    return py::make_tuple(ids_to_int64(p.ids), p.dists, ms);
  }

  py::dict stats() {
    vsg_server_stats st{};
    if (!s_ || vsg_server_get_stats(s_, &st) != VSG_OK) throw std::runtime_error("server is closed");
    py::dict d;
    d["requests"] = st.requests;
    d["queries"] = st.queries;
    d["batches"] = st.batches;
    d["admission_waits"] = st.admission_waits;
    d["peak_inflight_bytes"] = st.peak_inflight_bytes;
    d["max_batch_requests"] = st.max_batch_requests;
    d["shortlist_fallbacks"] = st.shortlist_fallbacks;
    return d;
  }

  void close() {
    if (!s_) return;
    {
      py::gil_scoped_release nogil;
      vsg_server_destroy(s_);
    }
    s_ = nullptr;
    pending_.clear();
  }

 private:
  struct Pending {
    FloatArray q;
    py::array_t<uint32_t, py::array::c_style | py::array::forcecast> allow;
    py::array_t<uint32_t> ids;
    py::array_t<float> dists;
  };
  py::object index_;
  size_t dim_ = 0;
  vsg_server* s_ = nullptr;
  std::unordered_map<uint64_t, Pending> pending_;
};

}  // namespace

PYBIND11_MODULE(_vecsearch, m) {
  m.doc() = "Exact k-NN on an AMD Instinct MI355X (HIP, gfx950)";

  py::enum_<Metric>(m, "Metric").value("L2", Metric::L2).value("InnerProduct", Metric::InnerProduct);

  py::class_<GpuBruteForceIndex>(m, "GpuBruteForceIndex")
      .def(py::init([](size_t dim, const std::string& metric, int device) {
             return std::make_unique<GpuBruteForceIndex>(dim, parse_metric(metric), device);
           }),
           py::arg("dim"), py::arg("metric") = "l2", py::arg("device") = 0)
      .def(
          "add",
          [](GpuBruteForceIndex& self, FloatArray x) {
            check_2d(x, self.dim());
            py::gil_scoped_release nogil;
            self.add(x.data(), x.shape(0));
          },
          py::arg("x"))
      .def(
          "search",
          [](GpuBruteForceIndex& self, FloatArray q, size_t k, const py::object& method_or_filter,
             const py::object& filter) {
            // The reference's third parameter is the method name (bindings.cpp:166-185); a mask
            // in that position is the filter, as in search(queries, k, filter).
            std::string method = "auto";
            py::object flt = filter;
            if (py::isinstance<py::str>(method_or_filter)) {
              method = method_or_filter.cast<std::string>();
            } else if (!method_or_filter.is_none()) {
              if (!filter.is_none())
                throw std::invalid_argument("search: a filter was given twice (third argument and filter=)");
              flt = method_or_filter;
            }
            const GpuMethod gm = parse_method(method);
            if (q.ndim() == 1) q = q.reshape({py::ssize_t(1), q.shape(0)});
            check_2d(q, self.dim());
            const size_t nq = q.shape(0);
            Mask mask = as_mask(flt, self.size());
            py::array_t<uint32_t> ids({nq, k});
            py::array_t<float> dists({nq, k});
            {
              py::gil_scoped_release nogil;
              if (mask.ptr)
                self.search(q.data(), nq, k, ids.mutable_data(), dists.mutable_data(), mask.ptr, gm);
              else
                self.search(q.data(), nq, k, ids.mutable_data(), dists.mutable_data(), gm);
            }
            return py::make_tuple(ids_to_int64(ids), dists);
          },
          py::arg("queries"), py::arg("k") = 10, py::arg("method") = "auto", py::arg("filter") = py::none(),
          "Exact top-k on the GPU, k <= 128. Returns (ids[int64, nq x k], distances[float32, nq x k]).\n"
          "method: 'auto', 'naive', 'skinny', 'tiled', 'bf16' or 'fp8' (the same results).\n"
          "filter: optional bool/uint8 mask over ids (len(self) entries); only allowed ids are returned.\n"
          "A mask passed as the third argument is taken as the filter: search(queries, k, filter).")
      .def(
          "search_batch",
          [](GpuBruteForceIndex& self, FloatArray q, py::array_t<uint32_t, py::array::c_style | py::array::forcecast> ks,
             py::array_t<uint64_t, py::array::c_style | py::array::forcecast> ranges,
             py::array_t<uint32_t, py::array::c_style | py::array::forcecast> allow_ids, py::object kmax_obj) {
            if (q.ndim() == 1) q = q.reshape({py::ssize_t(1), q.shape(0)});
            check_2d(q, self.dim());
            const size_t nq = q.shape(0);
            if (ks.ndim() != 1 || size_t(ks.shape(0)) != nq)
              throw std::invalid_argument("ks must be a 1-D array with one k per query");
            if (ranges.ndim() != 2 || size_t(ranges.shape(0)) != nq || ranges.shape(1) != 2)
              throw std::invalid_argument("ranges must have shape (nq, 2): [begin, end) into allow_ids");
            if (allow_ids.ndim() != 1)
              throw std::invalid_argument("allow_ids must be a 1-D array");
            const size_t n_ids = size_t(allow_ids.shape(0));
            size_t kmax = 0;
            for (size_t i = 0; i < nq; ++i) kmax = std::max<size_t>(kmax, ks.data()[i]);
            if (!kmax_obj.is_none()) kmax = std::max(kmax, kmax_obj.cast<size_t>());
            py::array_t<uint32_t> ids({nq, kmax});
            py::array_t<float> dists({nq, kmax});
            {
              py::gil_scoped_release nogil;
              self.search_lists(q.data(), nq, ks.data(), kmax, ranges.data(), allow_ids.data(),
                                n_ids, ids.mutable_data(), dists.mutable_data());
            }
            return py::make_tuple(ids_to_int64(ids), dists);
          },
          py::arg("queries"), py::arg("ks"), py::arg("ranges"), py::arg("allow_ids"),
          py::arg("kmax") = py::none(),
          "Batched filtered exact scan. Query i returns ks[i] results among ids\n"
          "allow_ids[ranges[i,0]:ranges[i,1]] (strictly increasing), or every row when\n"
          "ranges[i,0] == ALL_ROWS (unfiltered, same launch); queries with identical\n"
          "ranges share row reads. Returns (ids[int64, nq x kmax], distances[float32, nq x kmax]);\n"
          "columns past ks[i] are -1 / inf.")
      .def(
          "search_device",
          [](GpuBruteForceIndex& self, uintptr_t d_queries, size_t nq, size_t k, uintptr_t d_ids,
             uintptr_t d_dists, uintptr_t stream, const std::string& method, uintptr_t d_mask) {
            const GpuMethod gm = parse_method(method);
            int rc;
            {
              py::gil_scoped_release nogil;
              rc = vsg_search_device(self.handle(), reinterpret_cast<const float*>(d_queries), nq, k,
                                     static_cast<int>(gm), reinterpret_cast<const uint8_t*>(d_mask),
                                     reinterpret_cast<uint32_t*>(d_ids),
                                     reinterpret_cast<float*>(d_dists), reinterpret_cast<void*>(stream));
            }
            if (rc == VSG_ERR_INVALID_ARGUMENT) throw std::invalid_argument(vsg_last_error());
            if (rc != VSG_OK) throw std::runtime_error(vsg_last_error());
          },
          py::arg("d_queries"), py::arg("nq"), py::arg("k"), py::arg("d_ids"), py::arg("d_dists"),
          py::arg("stream") = 0, py::arg("method") = "auto", py::arg("d_mask") = 0,
          "Enqueue a search on device buffers (raw device pointers: queries nq x dim float32,\n"
          "ids nq x k uint32, dists nq x k float32) on `stream` (0 = the index's own stream).")
      .def(
          "kernel_timing",
          [](GpuBruteForceIndex& self, bool enable) {
            if (vsg_kernel_timing(self.handle(), enable ? 1 : 0) != VSG_OK)
              throw std::runtime_error(vsg_last_error());
          },
          py::arg("enable"))
      .def("kernel_times",
           [](GpuBruteForceIndex& self) {
             double ms[VSG_KERNEL_KINDS];
             uint64_t n[VSG_KERNEL_KINDS];
             if (vsg_kernel_times(self.handle(), ms, n) != VSG_OK) throw std::runtime_error(vsg_last_error());
             py::dict d;
             const char* names[VSG_KERNEL_KINDS] = {"scan", "merge", "mfma", "rescore", "sample", "lists"};
             for (int i = 0; i < VSG_KERNEL_KINDS; ++i) {
               d[(std::string(names[i]) + "_ms").c_str()] = ms[i];
               d[(std::string(names[i]) + "_launches").c_str()] = n[i];
             }
             return d;
           },
           "Accumulated kernel milliseconds and launch counts per kind since the last call (HIP events).")
      .def("short_stats",
           [](GpuBruteForceIndex& self) {
             uint64_t e = 0, r = 0, f = 0;
             if (vsg_short_stats(self.handle(), &e, &r, &f) != VSG_OK) throw std::runtime_error(vsg_last_error());
             py::dict d;
             d["emitted"] = e;
             d["rescored"] = r;
             d["fallbacks"] = f;
             return d;
           },
           "Shortlist counters of the last search that used it (emitted / rescored candidates are\n"
           "summed over queries only while kernel timing is on; fallbacks always).")
      .def(
          "set_param",
          [](GpuBruteForceIndex& self, const std::string& name, int64_t value) {
            if (vsg_set_param(self.handle(), name.c_str(), value) != VSG_OK)
              throw std::invalid_argument(vsg_last_error());
          },
          py::arg("name"), py::arg("value"), "Set a launch-shape parameter (see vsg.h).")
      .def(
          "get_param",
          [](GpuBruteForceIndex& self, const std::string& name) {
            int64_t v = 0;
            if (vsg_get_param(self.handle(), name.c_str(), &v) != VSG_OK)
              throw std::invalid_argument(vsg_last_error());
            return v;
          },
          py::arg("name"), "Read a launch-shape parameter or tuner state (see vsg.h).")
      .def("synchronize", [](GpuBruteForceIndex& self) {
        py::gil_scoped_release nogil;
        if (vsg_synchronize(self.handle()) != VSG_OK) throw std::runtime_error(vsg_last_error());
      })
      .def_property_readonly("dim", &GpuBruteForceIndex::dim)
      .def_property_readonly("metric", [](const GpuBruteForceIndex& self) {
        return self.metric() == Metric::L2 ? "l2" : "ip";
      })
      .def("__len__", &GpuBruteForceIndex::size)
      .def(
          "add_hnsw_file",
          [](GpuBruteForceIndex& self, const std::string& path) {
            size_t rows = 0;
            int rc;
            {
              py::gil_scoped_release nogil;
              rc = vsg_add_hnsw_file(self.handle(), path.c_str(), &rows);
            }
            if (rc == VSG_ERR_INVALID_ARGUMENT) throw std::invalid_argument(vsg_last_error());
            if (rc != VSG_OK) throw std::runtime_error(vsg_last_error());
            return rows;
          },
          py::arg("path"),
          "Append the vector block of a saved HNSWIndex file (hnsw.cpp:375-389) to this index;\n"
          "returns the number of rows added. The file's dim and metric must match.")
      .def_static(
          "from_hnsw_file",
          [](const std::string& path, int device) {
            size_t dim = 0, n = 0;
            int metric = 0;
            if (vsg_hnsw_file_info(path.c_str(), &dim, &n, &metric) != VSG_OK)
              throw std::invalid_argument(vsg_last_error());
            auto idx = std::make_unique<GpuBruteForceIndex>(dim, metric == 0 ? Metric::L2 : Metric::InnerProduct,
                                                            device);
            int rc;
            {
              py::gil_scoped_release nogil;
              rc = vsg_add_hnsw_file(idx->handle(), path.c_str(), nullptr);
            }
            if (rc == VSG_ERR_INVALID_ARGUMENT) throw std::invalid_argument(vsg_last_error());
            if (rc != VSG_OK) throw std::runtime_error(vsg_last_error());
            return idx;
          },
          py::arg("path"), py::arg("device") = 0,
          "A GPU index holding the vectors of a saved HNSWIndex file (same ids, dim and metric).")
      .def(
          "search_exact_below",
          [](GpuBruteForceIndex& self, FloatArray q, size_t k, const py::object& filter,
             size_t exact_below) -> py::object {
            if (q.ndim() == 1) q = q.reshape({py::ssize_t(1), q.shape(0)});
            check_2d(q, self.dim());
            const size_t nq = q.shape(0);
            Mask mask = as_mask(filter, self.size());
            py::array_t<uint32_t> ids({nq, k});
            py::array_t<float> dists({nq, k});
            int handled = 0, rc;
            {
              py::gil_scoped_release nogil;
              rc = vsg_search_exact_below(self.handle(), q.data(), nq, k, mask.ptr, exact_below,
                                          ids.mutable_data(), dists.mutable_data(), &handled);
            }
            if (rc == VSG_ERR_INVALID_ARGUMENT) throw std::invalid_argument(vsg_last_error());
            if (rc != VSG_OK) throw std::runtime_error(vsg_last_error());
            if (!handled) return py::none();
            return py::make_tuple(ids_to_int64(ids), dists);
          },
          py::arg("queries"), py::arg("k") = 10, py::arg("filter") = py::none(), py::arg("exact_below") = 0,
          "HNSWIndex.search's filter dispatch (hnsw.cpp:322-327): with a filter allowing at most\n"
          "exact_below ids, the exact scan over them (ids, distances); otherwise None (graph search).");

  py::class_<PyGraph>(m, "SearchGraph")
      .def(py::init<py::object, uintptr_t, size_t, size_t, uintptr_t, uintptr_t, const std::string&>(),
           py::arg("index"), py::arg("d_queries"), py::arg("nq"), py::arg("k"), py::arg("d_ids"),
           py::arg("d_dists"), py::arg("method") = "skinny",
           "Capture one search of nq device-resident queries into a HIP graph (vsg_graph_create): the\n"
           "streaming scan, or the certified shortlist (tiled / bf16 / fp8 / auto from auto_short_batch\n"
           "queries) with its exact fallback planned on the device: the path the eager search takes.")
      .def_static(
          "lists",
          [](py::object index, uintptr_t d_q, py::array_t<uint32_t, py::array::c_style | py::array::forcecast> ks,
             py::array_t<uint64_t, py::array::c_style | py::array::forcecast> ranges,
             py::array_t<uint32_t, py::array::c_style | py::array::forcecast> allow_ids, uintptr_t d_ids,
             uintptr_t d_dists, size_t kmax) {
            return new PyGraph(std::move(index), d_q, ks, ranges, allow_ids, kmax, d_ids, d_dists);
          },
          py::arg("index"), py::arg("d_queries"), py::arg("ks"), py::arg("ranges"), py::arg("allow_ids"),
          py::arg("d_ids"), py::arg("d_dists"), py::arg("kmax") = 0,
          "Capture one persistent launch of the batched filtered scan (search_batch's ks, ranges and\n"
          "allow_ids, fixed at capture) over len(ks) device-resident queries; results are\n"
          "len(ks) x kmax (kmax >= max(ks)) in d_ids (uint32) / d_dists (vsg_graph_create_lists).")
      .def("launch", &PyGraph::launch, py::arg("stream") = 0, "Replay the captured search on a HIP stream.")
      .def("close", &PyGraph::close);

  py::class_<PyServer>(m, "Server")
      .def(py::init<py::object, size_t, size_t, int, const std::string&, bool>(), py::arg("index"),
           py::arg("max_batch_queries") = 4096, py::arg("workspace_bytes") = size_t(4) << 30,
           py::arg("slots") = 2, py::arg("method") = "auto", py::arg("tune") = true,
           "Continuous-batching request loop over an index (see vsg.h, vsg_server_*).")
      .def("submit", &PyServer::submit, py::arg("queries"), py::arg("k") = 10,
           py::arg("allow_ids") = py::none(),
           "Queue a request; returns its ticket. allow_ids: increasing ids the request may return.")
      .def("wait", &PyServer::wait, py::arg("ticket"),
           "Block until the request is done: (ids[int64, nq x k], distances[float32], latency_ms).")
      .def("stats", &PyServer::stats)
      .def("close", &PyServer::close)
      .def("__enter__", [](PyServer& s) -> PyServer& { return s; })
      .def("__exit__", [](PyServer& s, py::object, py::object, py::object) { s.close(); });

  m.attr("ALL_ROWS") = py::int_(static_cast<unsigned long long>(VSG_ALL_ROWS));
  m.def(
      "hnsw_file_info",
      [](const std::string& path) {
        size_t dim = 0, n = 0;
        int metric = 0;
        if (vsg_hnsw_file_info(path.c_str(), &dim, &n, &metric) != VSG_OK)
          throw std::invalid_argument(vsg_last_error());
        return py::make_tuple(dim, n, metric == 0 ? "l2" : "ip");
      },
      py::arg("path"), "(dim, n, metric) from the header of a saved HNSWIndex file.");
  m.def("cuda_available", [] { return cuda_available(); },
        "True if a GPU is visible (API-compatible name; the runtime is HIP).");
  m.def("hip_available", [] { return hip_available(); }, "True if a HIP GPU is visible.");
}

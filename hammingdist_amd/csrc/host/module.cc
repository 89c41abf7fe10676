// pybind11 module of hammingdist_amd: the reference's Python API
// (python/hammingdist.cc:28-133) with the same names, arguments, defaults,
// return types and error messages, over the HIP engine.
#include <memory>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "hamming.hh"
#include "synth.hh"

namespace py = pybind11;

namespace {

// Hands a vector to numpy without a copy (python/hammingdist.cc:13-24).
template <typename Seq> py::array_t<typename Seq::value_type> as_pyarray(Seq &&seq) {
  auto size = seq.size();
  auto data = seq.data();
  auto owner = std::make_unique<Seq>(std::move(seq));
  auto capsule = py::capsule(owner.get(), [](void *p) { delete reinterpret_cast<Seq *>(p); });
  owner.release();
  return py::array(size, data, capsule);
}

template <typename T>
void bind_dataset(py::module_ &m, const char *name, int default_threshold) {
  using DS = hdhost::DataSet<T>;
  py::class_<DS>(m, name)
      .def("dump", &DS::dump, "Dump distances matrix in csv format")
      .def("dump_lower_triangular", &DS::dump_lower_triangular,
           "Dump distances matrix in lower triangular format (comma-delimited, "
           "row-major)")
      .def("dump_sparse", &DS::dump_sparse, py::arg("filename"),
           py::arg("threshold") = default_threshold,
           "Dump distances matrix in sparse format excluding any distances "
           "above threshold")
      .def("dump_sequence_indices", &DS::dump_sequence_indices,
           "Dump row index in distances matrix for each input sequence")
      .def("__getitem__",
           [](const DS &self, const std::array<std::size_t, 2> &ij) {
             return self.get(ij[0], ij[1]);
           })
      .def_readonly("_distances", &DS::result)
      .def_readonly("nsamples", &DS::nsamples)
      .def_property_readonly("lt_array", [](const DS &self) {
        return py::array(self.result.size(), self.result.data());
      });
}

hdhost::SynthSpec spec_from(std::uint64_t n, std::uint64_t L, std::uint64_t seed,
                            double p_sub, std::uint32_t clades, double p_clade,
                            double gap_frac, double gap_mean, double p_n, double p_x,
                            double dup_frac) {
  hdhost::SynthSpec s;
  s.n = n;
  s.L = L;
  s.seed = seed;
  s.p_sub = p_sub;
  s.clades = clades;
  s.p_clade = p_clade;
  s.gap_frac = gap_frac;
  s.gap_mean = gap_mean;
  s.p_n = p_n;
  s.p_x = p_x;
  s.dup_frac = dup_frac;
  return s;
}

#define HD_SYNTH_ARGS                                                              \
  py::arg("n"), py::arg("L"), py::arg("seed"), py::arg("p_sub") = 0.0,              \
      py::arg("clades") = 0, py::arg("p_clade") = 0.0, py::arg("gap_frac") = 0.0,   \
      py::arg("gap_mean") = 20.0, py::arg("p_n") = 0.0, py::arg("p_x") = 0.0,       \
      py::arg("dup_frac") = 0.0

} // namespace

// The reference's module name (python/hammingdist.cc:28), so that callers
// written against it `import hammingdist` unchanged (the top-level
// hammingdist package re-exports this module).
PYBIND11_MODULE(hammingdist, m) {
  using namespace hdhost;
  m.doc() = "Small tool to calculate Hamming distances between gene sequences "
            "(MI355X HIP engine)";

  bind_dataset<std::uint8_t>(m, "DataSet", 255);
  bind_dataset<std::uint16_t>(m, "DataSetLarge", 65535);

  m.def("from_stringlist", &from_stringlist, "Creates a dataset from a list of strings");
  m.def("from_csv", &from_csv,
        "Creates a dataset by reading already computed distances from csv "
        "(full matrix expected)");
  m.def("from_fasta", &from_fasta<std::uint8_t>, py::arg("filename"),
        py::arg("include_x") = false, py::arg("remove_duplicates") = false,
        py::arg("n") = 0, py::arg("use_gpu") = false, py::arg("max_distance") = 255,
        "Creates a dataset by reading from a fasta file (assuming all "
        "sequences have equal length). Maximum value of an element in the "
        "distances matrix: max_distance or 255, whichever is lower."
        "Distances that would have been larger than "
        "this value instead saturate at this value - to support genomes with "
        "larger distances than this see `from_fasta_large` instead.");
  m.def("from_fasta_large", &from_fasta<std::uint16_t>, py::arg("filename"),
        py::arg("include_x") = false, py::arg("remove_duplicates") = false,
        py::arg("n") = 0, py::arg("use_gpu") = false, py::arg("max_distance") = 65535,
        "Creates a dataset by reading from a fasta file (assuming all "
        "sequences have equal length). Maximum value of an element in the "
        "distances matrix: max_distance or 65535, whichever is lower");
  m.def("from_fasta_to_lower_triangular", &from_fasta_to_lower_triangular,
        py::arg("fasta_filename"), py::arg("output_filename"),
        py::arg("remove_duplicates") = false, py::arg("n") = 0, py::arg("use_gpu") = true,
        py::arg("max_distance") = 65535,
        "Construct lower triangular distances matrix output file from the "
        "fasta file, requires a GPU. Maximum value of an element in "
        "the distances matrix: max_distance or 65535, whichever is lower");
  m.def("from_lower_triangular", &from_lower_triangular<std::uint8_t>,
        "Creates a dataset by reading already computed distances from lower "
        "triangular format. Maximum value of an element in the distances "
        "matrix: 255.");
  m.def("from_lower_triangular_large", &from_lower_triangular<std::uint16_t>,
        "Creates a dataset by reading already computed distances from lower "
        "triangular format. Maximum value of an element in the distances "
        "matrix: 65535.");
  m.def("distance", &distance, py::arg("seq0"), py::arg("seq1"),
        py::arg("include_x") = false, "Calculate the distance between seq0 and seq1");
  m.def(
      "fasta_reference_distances",
      [](const std::string &reference_sequence, const std::string &fasta_file,
         bool include_x) {
        return as_pyarray(fasta_reference_distances(reference_sequence, fasta_file, include_x));
      },
      py::arg("reference_sequence"), py::arg("fasta_file"), py::arg("include_x") = false,
      "Calculates the distance of each sequence in the fasta file from the "
      "supplied reference sequence");
  m.def(
      "fasta_sequence_indices",
      [](const std::string &fasta_file, std::size_t n) {
        return as_pyarray(fasta_sequence_indices(fasta_file, n));
      },
      py::arg("fasta_file"), py::arg("n") = 0,
      "Returns the same output as dump_sequence_indices() but without "
      "constructing the distances matrix."
      "For each genome in the input fasta file it gives the index of the "
      "corresponding row in the distances matrix which excludes duplicates");
  m.def("cuda_gpu_available", &gpu_available,
        "True if a GPU (here: a HIP device) is available");

  // ---- engine extras (no counterpart in the reference) -------------------
  m.def("_bytes_per_seq", &bytes_per_seq, py::arg("L"));
  m.def(
      "_read_fasta",
      [](const std::string &filename, bool remove_duplicates, std::size_t n) {
        FastaData fa = read_fasta(filename, remove_duplicates, n);
        py::list data;
        for (const auto &s : fa.data) data.append(py::bytes(s));
        return py::make_tuple(data, fa.sequence_indices);
      },
      py::arg("filename"), py::arg("remove_duplicates") = false, py::arg("n") = 0,
      "(sequences as bytes, sequence indices) as read_fasta parses the file");
  m.def(
      "_write_row_bands",
      [](const std::string &filename, py::array cells, std::vector<std::uint64_t> bounds,
         std::uint32_t max_value) {
        // One RowWriter fed band by band, as the streaming triangle is.
        const py::buffer_info info = cells.request();
        if (info.ndim != 1 || (info.itemsize != 1 && info.itemsize != 2))
          throw std::runtime_error("cells must be a flat uint8 or uint16 array");
        RowWriter w(filename, false);
        for (std::size_t b = 0; b + 1 < bounds.size(); ++b) {
          const std::uint64_t r0 = bounds[b], c0 = r0 < 1 ? 0 : r0 * (r0 - 1) / 2;
          if (info.itemsize == 1)
            w.write(static_cast<const std::uint8_t *>(info.ptr) + c0, r0, bounds[b + 1],
                    max_value);
          else
            w.write(static_cast<const std::uint16_t *>(info.ptr) + c0, r0, bounds[b + 1],
                    max_value);
        }
      },
      py::arg("filename"), py::arg("cells"), py::arg("bounds"), py::arg("max_value"),
      "Writes the triangle's rows [bounds[b], bounds[b+1]) band after band");
  m.def(
      "_pack",
      [](const std::vector<std::string> &data, bool include_x) {
        const std::uint64_t L = data.empty() ? 0 : data[0].size();
        for (const auto &s : data)
          if (s.size() != L)
            throw std::runtime_error("Error: Sequences do not all have the same length");
        return as_pyarray(pack_all(data, include_x, bytes_per_seq(L)));
      },
      py::arg("data"), py::arg("include_x") = false,
      "Packed rows (reference dense format) of equal-length sequences, flat uint8");
  m.def(
      "_synth_chars",
      [](std::uint64_t n, std::uint64_t L, std::uint64_t seed, double p_sub,
         std::uint32_t clades, double p_clade, double gap_frac, double gap_mean,
         double p_n, double p_x, double dup_frac) {
        auto v = synth_chars(
            spec_from(n, L, seed, p_sub, clades, p_clade, gap_frac, gap_mean, p_n, p_x, dup_frac));
        std::vector<std::uint8_t> u(v.begin(), v.end());
        return as_pyarray(std::move(u));
      },
      HD_SYNTH_ARGS, "n*L synthetic sequence chars (uint8)");
  m.def(
      "_synth_packed",
      [](std::uint64_t n, std::uint64_t L, std::uint64_t seed, double p_sub,
         std::uint32_t clades, double p_clade, double gap_frac, double gap_mean,
         double p_n, double p_x, double dup_frac, bool include_x) {
        py::gil_scoped_release release;
        auto v = synth_packed(
            spec_from(n, L, seed, p_sub, clades, p_clade, gap_frac, gap_mean, p_n, p_x, dup_frac),
            include_x);
        py::gil_scoped_acquire acquire;
        return as_pyarray(std::move(v));
      },
      HD_SYNTH_ARGS, py::arg("include_x") = false,
      "Synthetic sequences packed to bytes_per_seq(L) each (uint8, flat)");
  m.def(
      "_synth_fasta",
      [](const std::string &path, std::uint64_t n, std::uint64_t L, std::uint64_t seed,
         double p_sub, std::uint32_t clades, double p_clade, double gap_frac,
         double gap_mean, double p_n, double p_x, double dup_frac) {
        synth_fasta(spec_from(n, L, seed, p_sub, clades, p_clade, gap_frac, gap_mean, p_n, p_x,
                              dup_frac),
                    path);
      },
      py::arg("path"), HD_SYNTH_ARGS, "Writes synthetic sequences as a FASTA file");
  m.def(
      "_synth_mutant",
      [](std::uint64_t L, std::uint64_t seed, double p_sub, double p_n, double p_x,
         std::uint64_t stream) {
        hdhost::SynthSpec s;
        s.L = L;
        s.seed = seed;
        s.p_n = p_n;
        s.p_x = p_x;
        return synth_mutant_of_root(s, p_sub, stream);
      },
      py::arg("L"), py::arg("seed"), py::arg("p_sub"), py::arg("p_n") = 0.0,
      py::arg("p_x") = 0.0, py::arg("stream") = 0,
      "A mutant of the synthetic root sequence of (L, seed)");
}

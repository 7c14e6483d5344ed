// Model3D (model.cc:10-51): Born-model the shots of data= from the image
// (reflectivity) image= on the grid of velocity=, with the wavelet of
// wavelet=.  data= must exist: its axes give the time axis, the receiver
// grid and the shot grid; its values are replaced.
//   Model3D velocity=v.json wavelet=w.json image=i.json data=d.json
//           [src_depth= rec_depth= nbound= scatter=0|1 storage=f32|f16]
#include <cstdio>
#include <exception>

#include "cli_common.h"

int main(int argc, char **argv) {
  try {
    born_io::Params par(argc, argv);
    born_cli::Inputs in = born_cli::read_inputs(par);
    const born_io::Cube image = born_io::read_cube(par.str("image"));
    born_op *op = born_cli::create_op(par, in);
    long long model_n = 0, data_n = 0;
    born_cli::check(op, born_op_sizes(op, &model_n, &data_n, nullptr, nullptr, nullptr),
                    "sizes");
    if ((long long)image.vals.size() != model_n)
      throw std::runtime_error("image holds " + std::to_string(image.vals.size()) +
                               " cells, the velocity grid " + std::to_string(model_n));
    born_io::Cube data;
    data.axes = in.data_axes;
    data.vals.assign((size_t)data_n, 0.f);
    born_cli::check(op, born_op_forward(op, 0, image.vals.data(), data.vals.data()), "forward");
    born_op_destroy(op);
    born_io::write_cube(par.str("data"), data);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "Model3D: %s\n", e.what());
    return 1;
  }
  return 0;
}

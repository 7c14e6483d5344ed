// RTM3D (rtm.cc:10-41): migrate the shots of data= into image= on the grid
// of velocity=, with the wavelet of wavelet=.
//   RTM3D velocity=v.json wavelet=w.json data=d.json image=i.json
//         [src_depth= rec_depth= nbound= mode=0|1 scatter=0|1 storage=f32|f16]
// mode 0 is the reference's RTM (random-boundary background rebuilt by time
// reversal), mode 1 the exact transpose of Model3D.
#include <cstdio>
#include <exception>

#include "cli_common.h"

int main(int argc, char **argv) {
  try {
    born_io::Params par(argc, argv);
    born_cli::Inputs in = born_cli::read_inputs(par);
    const born_io::Cube data = born_io::read_cube(par.str("data"));
    born_op *op = born_cli::create_op(par, in);
    long long model_n = 0, data_n = 0;
    born_cli::check(op, born_op_sizes(op, &model_n, &data_n, nullptr, nullptr, nullptr),
                    "sizes");
    if ((long long)data.vals.size() != data_n)
      throw std::runtime_error("data holds " + std::to_string(data.vals.size()) +
                               " samples, the geometry " + std::to_string(data_n));
    born_io::Cube image;
    image.axes = in.vel.axes;
    image.vals.assign((size_t)model_n, 0.f);
    born_cli::check(op, born_op_adjoint(op, 0, image.vals.data(), data.vals.data()), "adjoint");
    born_op_destroy(op);
    born_io::write_cube(par.str("image"), image);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "RTM3D: %s\n", e.what());
    return 1;
  }
  return 0;
}

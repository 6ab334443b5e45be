! Links a Fortran program against libhochunk3d_amd.so through hc_iface and
! checks the ABI version and the sizes of the bind(C) types (no GPU call).
program hc_abi_check
  use hc_iface
  implicit none
  type(hc_dims) :: d
  type(hc_sources) :: s
  type(hc_tallies) :: t
  type(hc_density_model) :: m
  print '(A,I0)', 'abi ', hc_abi_version()
  print '(A,I0,1X,I0,1X,I0,1X,I0)', 'sizes ', storage_size(d) / 8, storage_size(s) / 8, &
       storage_size(t) / 8, storage_size(m) / 8
end program hc_abi_check

! Test-only harness around the reference's streamer/fosito modules
! (m_streamer, m_lookup_table, m_transport_data, m_morton), compiled by
! oracle/Makefile into oracle/_ref/. Exposes C-callable entry points for the
! Koren limiter, the uniform lookup table, the transport-data text reader and
! Morton ordering. Nothing in the product links or loads this.

module ref_streamer
  use iso_c_binding
  use m_lookup_table
  use m_morton
  use m_streamer, only: koren_mlim
  use omp_lib, only: omp_set_num_threads
  use m_vtk, only: vtk_t, vtk_ini_xml, vtk_dat_xml, vtk_geo_xml, vtk_con_xml, &
       vtk_var_r8_xml, vtk_geo_xml_close, vtk_end_xml
  implicit none
  private

  integer, parameter :: dp = kind(0.0d0)
  type(LT_table_t), save :: tbl

contains

  function h_koren(a, b) result(r) bind(C, name="ref_koren_mlim")
    real(c_double), value :: a, b
    real(c_double)        :: r
    r = koren_mlim(a, b)
  end function h_koren

  subroutine h_lt_create(x_min, x_max, n_rows, n_cols) bind(C, name="ref_lt_create")
    real(c_double), value :: x_min, x_max
    integer(c_int), value :: n_rows, n_cols
    tbl = LT_create(x_min, x_max, n_rows, n_cols)
  end subroutine h_lt_create

  subroutine h_lt_set_col(col, n, xx, yy) bind(C, name="ref_lt_set_col")
    integer(c_int), value :: col, n
    real(c_double)        :: xx(n), yy(n)
    call LT_set_col(tbl, col, xx, yy)
  end subroutine h_lt_set_col

  function h_lt_get_col(col, x) result(y) bind(C, name="ref_lt_get_col")
    integer(c_int), value :: col
    real(c_double), value :: x
    real(c_double)        :: y
    y = LT_get_col(tbl, col, x)
  end function h_lt_get_col

  subroutine h_lt_get_data(n_rows, n_cols, x_data, cols_rows) bind(C, name="ref_lt_get_data")
    integer(c_int), value :: n_rows, n_cols
    real(c_double)        :: x_data(n_rows), cols_rows(n_cols, n_rows)
    call LT_get_data(tbl, x_data, cols_rows)
  end subroutine h_lt_get_data

  subroutine h_lin_interp(n, xx, yy, x, y) bind(C, name="ref_lin_interp_list")
    integer(c_int), value       :: n
    real(c_double)              :: xx(n), yy(n)
    real(c_double), value       :: x
    real(c_double), intent(out) :: y
    call LT_lin_interp_list(xx, yy, x, y)
  end subroutine h_lin_interp

  ! Reads one transport-data entry; n_out is the number of rows found
  subroutine h_td_read(fname, gname, dname, max_n, n_out, xx, yy) &
       bind(C, name="ref_td_read")
    use m_transport_data
    character(kind=c_char) :: fname(*), gname(*), dname(*)
    integer(c_int), value       :: max_n
    integer(c_int), intent(out) :: n_out
    real(c_double)              :: xx(max_n), yy(max_n)
    real(dp), allocatable       :: x_data(:), y_data(:)
    call TD_get_td_from_file(c2f(fname), c2f(gname), c2f(dname), x_data, y_data)
    n_out = min(size(x_data), max_n)
    xx(1:n_out) = x_data(1:n_out)
    yy(1:n_out) = y_data(1:n_out)
  end subroutine h_td_read

  function c2f(s) result(f)
    character(kind=c_char) :: s(*)
    character(len=:), allocatable :: f
    integer :: n, i
    n = 0
    do while (s(n+1) /= c_null_char)
       n = n + 1
    end do
    allocate(character(len=n) :: f)
    do i = 1, n
       f(i:i) = s(i)
    end do
  end function c2f

  ! the compiler's NORM2 intrinsic, which the update's source term uses
  function h_norm2(n, x) result(r) bind(C, name="ref_norm2")
    integer(c_int), value :: n
    real(c_double)        :: x(n)
    real(c_double)        :: r
    r = norm2(x)
  end function h_norm2

  function h_morton2(i, j) result(m) bind(C, name="ref_morton_from_ix2")
    integer(c_int), value :: i, j
    integer(c_int64_t)    :: m
    m = morton_from_ix2([i, j])
  end function h_morton2

  function h_morton3(i, j, k) result(m) bind(C, name="ref_morton_from_ix3")
    integer(c_int), value :: i, j, k
    integer(c_int64_t)    :: m
    m = morton_from_ix3([i, j, k])
  end function h_morton3

  subroutine h_morton_rank(n, mortons, ranks) bind(C, name="ref_morton_rank")
    integer(c_int), value :: n
    integer(c_int64_t)    :: mortons(n)
    integer(c_int)        :: ranks(n)
    integer(morton_k)     :: m(n)
    integer               :: r(n)
    m = mortons
    call morton_rank(m, r)
    ranks = r
  end subroutine h_morton_rank

  ! A VTK unstructured grid written by the reference's m_vtk with the calls of
  ! a$D_write_vtk (afivo/src/m_aX_io.f90:190-203); names are 100 characters
  ! each, blank padded
  subroutine h_vtk_write(fname, n_nodes, n_cells, n_dim, coords, connects, offsets, &
       cell_types, n_vars, names, vars, cycl, time) bind(C, name="ref_vtk_write")
    character(kind=c_char), intent(in) :: fname(*)
    integer(c_int), value              :: n_nodes, n_cells, n_dim, n_vars, cycl
    real(c_double), intent(in)         :: coords(n_dim * n_nodes), vars(n_cells, n_vars)
    integer(c_int), intent(in)         :: connects(2**n_dim * n_cells), offsets(n_cells)
    integer(c_int), intent(in)         :: cell_types(n_cells)
    character(kind=c_char), intent(in) :: names(100 * n_vars)
    real(c_double), value              :: time
    type(vtk_t)                        :: vtkf
    character(len=400)                 :: f
    character(len=100)                 :: nm
    integer                            :: i, n

    f = ''
    do i = 1, 400
       if (fname(i) == c_null_char) exit
       f(i:i) = fname(i)
    end do
    call vtk_ini_xml(vtkf, trim(f), 'UnstructuredGrid')
    call vtk_dat_xml(vtkf, "UnstructuredGrid", .true.)
    call vtk_geo_xml(vtkf, coords, n_nodes, n_cells, n_dim, cycl, time)
    call vtk_con_xml(vtkf, connects, offsets, cell_types, n_cells)
    call vtk_dat_xml(vtkf, "CellData", .true.)
    do n = 1, n_vars
       do i = 1, 100
          nm(i:i) = names(100 * (n - 1) + i)
       end do
       call vtk_var_r8_xml(vtkf, trim(nm), vars(:, n), n_cells)
    end do
    call vtk_dat_xml(vtkf, "CellData", .false.)
    call vtk_geo_xml_close(vtkf)
    call vtk_dat_xml(vtkf, "UnstructuredGrid", .false.)
    call vtk_end_xml(vtkf)
  end subroutine h_vtk_write

  ! OpenMP threads of the reference routines called after this
  subroutine h_set_threads(n) bind(C, name="ref_set_threads")
    integer(c_int), value :: n
    call omp_set_num_threads(n)
  end subroutine h_set_threads

end module ref_streamer
